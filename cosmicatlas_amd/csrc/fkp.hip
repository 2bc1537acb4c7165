// FKP estimator with a random catalogue (gfx950).
//
// Per catalogue (Catalog::get_interpolated_density_field, Source/Catalog.cpp:
// 14153-14171): total weight w = w_user * w_fkp, w_fkp = 1/(1 + nbar P_est);
// W = sum w, S = sum w^2, normal = sum nbar w^2, with nbar per object
// (Halo.mean_density, :14150-14152) or one constant.
// FftwFunctions::get_parameters_estimator random branch
// (Source/FftwFunctions.cpp:180-200): alpha = W_g/W_r, normal_power =
// alpha normal_r, shot noise = (alpha + 1) S_r/normal_r, window
// normalisation normal_power/alpha^2.  get_fluctuation (:275-281):
// delta = rho_g - alpha rho_r, built on one mesh by depositing the randoms with
// weight -alpha w.  The shells are those of the fused path.
#include <cmath>
#include <vector>

#include "cm_internal.hpp"

namespace cm {

// The total weight of object i (Source/Catalog.cpp:14142-14160): its user
// weight (1 without weights) times, when fkp, 1/(1 + P_est nbar) with the
// object's own mean density nb[i], or nbar0 for all objects when nb is NULL.
struct WeightRule {
    const float *w;
    const float *nb;
    double nbar0, P_est;
    int fkp;
};

__device__ inline double total_weight(const WeightRule &r, size_t i, double &nb) {
    nb = r.nb ? (double)r.nb[i] : r.nbar0;
    double v = r.w ? (double)r.w[i] : 1.0;
    if (r.fkp) v *= 1.0 / (1.0 + r.P_est * nb);
    return v;
}

// FP64 sums of w, w^2 and nbar w^2, one partial triple per workgroup.
__global__ void __launch_bounds__(256) k_wsums(WeightRule r, size_t n, double *__restrict__ part) {
    __shared__ double r1[256], r2[256], r3[256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double nb;
        const double v = total_weight(r, i, nb);
        a += v;
        b += v * v;
        c += nb * v * v;
    }
    r1[threadIdx.x] = a;
    r2[threadIdx.x] = b;
    r3[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
            r3[threadIdx.x] += r3[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = r1[0];
        part[3 * blockIdx.x + 1] = r2[0];
        part[3 * blockIdx.x + 2] = r3[0];
    }
}

// deposit weights out[i] = s * (total weight of object i)
__global__ void __launch_bounds__(256) k_wscale(WeightRule r, size_t n, double s, float *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double nb;
        out[i] = (float)(s * total_weight(r, i, nb));
    }
}

static unsigned grid_of(size_t n, size_t cap) {
    size_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

struct WeightSums {
    double W = 0, S = 0, normal = 0;  // sum w, sum w^2, sum nbar w^2
};

static cm_status weight_sums(cm_ctx *ctx, const WeightRule &r, size_t n, WeightSums *out) {
    *out = WeightSums{};
    if (!n) return CM_OK;
    const unsigned g = grid_of(n, 1024);
    void *part;
    cm_status s;
    if ((s = scratch(ctx, "fkp.part", sizeof(double) * 3 * g, &part)) != CM_OK) return s;
    if ((s = launch(ctx, "fkp_wsums", k_wsums, dim3(g), dim3(256), 0, r, n, (double *)part)) != CM_OK) return s;
    std::vector<double> h(3 * (size_t)g);
    CM_HIP(ctx, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (unsigned i = 0; i < g; ++i) {
        out->W += h[3 * i];
        out->S += h[3 * i + 1];
        out->normal += h[3 * i + 2];
    }
    return CM_OK;
}

// deposit weights s * w_i; NULL when they are all 1
static cm_status scaled_weights(cm_ctx *ctx, const char *name, const WeightRule &r, size_t n, double s,
                                const float **out) {
    *out = nullptr;
    if (!r.w && !r.fkp && s == 1.0) return CM_OK;
    void *buf;
    cm_status st;
    if ((st = scratch(ctx, name, sizeof(float) * (n ? n : 1), &buf)) != CM_OK) return st;
    if (n && (st = launch(ctx, "fkp_wscale", k_wscale, dim3(grid_of(n, 16384)), dim3(256), 0, r, n, s,
                          (float *)buf)) != CM_OK)
        return st;
    *out = (const float *)buf;
    return CM_OK;
}

// mesh -> shells through the fused path, host sums, counts and the FP64 mesh sum
static cm_status shells_of_mesh(cm_ctx *ctx, const cm_measure_opts *o, void *mesh, double mean,
                                std::vector<double> &hs, std::vector<unsigned long long> &hc,
                                bool window_set = false, double *mesh_sum = nullptr) {
    FusedOut dev, h;
    cm_status s;
    if ((s = fused_measure_mesh(ctx, o, mesh, nullptr, mean, nullptr, &dev, window_set)) != CM_OK) return s;
    if ((s = fused_readback(ctx, dev, o->nshells, &h)) != CM_OK) return s;
    hs.assign(h.sums, h.sums + 4 * (size_t)o->nshells);
    hc.assign(h.cnts, h.cnts + (size_t)o->nshells);
    if (mesh_sum) *mesh_sum = *h.nobj;
    return CM_OK;
}

// A random catalogue has no spatial order: its particles are deposited through
// (at least) four x-slab groups, so that the binning records of one group are
// written while that group's part of the record buffer stays in the caches
// (one host synchronisation for the group sizes).
static cm_status assign_randoms(cm_ctx *ctx, const cm_mas_desc *md, const void *rpos, const float *wr, size_t nr,
                                void *mesh, int accumulate) {
    const int saved = ctx->tune.mas_groups;
    ctx->tune.mas_groups = 4;
    const cm_status s = mesh_assign_impl(ctx, md, rpos, wr, nr, mesh, CM_F32, accumulate);
    ctx->tune.mas_groups = saved;
    return s;
}

static cm_status measure_fkp_impl(cm_ctx *ctx, const cm_measure_opts *o, const cm_fkp_opts *f, const void *pos,
                                  const float *w, size_t n, const void *rpos, const float *rw, size_t nr,
                                  double *pk0, double *pk2, double *pk4, int64_t *nmodes, double *pk_window,
                                  cm_measure_info *info) {
    const int N = o->mas.N;
    const size_t n3 = (size_t)N * N * N;
    const double L = o->mas.L, V = L * L * L;
    cm_status s;
    // the constant mean density: given, or the objects' user weight over the volume
    WeightRule rg{w, nullptr, 0.0, 0.0, 0};
    WeightSums g0, g, r;
    if ((s = weight_sums(ctx, rg, n, &g0)) != CM_OK) return s;
    const double nbar = f->nbar > 0 ? f->nbar : g0.W / V;
    rg = WeightRule{w, (const float *)f->nbar_g, nbar, f->P_est, f->fkp_weight};
    const WeightRule rr{rw, (const float *)f->nbar_r, nbar, f->P_est, f->fkp_weight};
    if (rg.fkp) {
        if ((s = weight_sums(ctx, rg, n, &g)) != CM_OK) return s;
    } else {
        g = g0;
    }
    if ((s = weight_sums(ctx, rr, nr, &r)) != CM_OK) return s;
    if (!(g.W > 0) || !(r.W > 0)) return fail(ctx, CM_ERR_INVALID, "cm_measure_fkp: empty catalogue");
    if (!(r.normal > 0)) return fail(ctx, CM_ERR_INVALID, "cm_measure_fkp: the randoms' sum of nbar w^2 is not positive");
    const double alpha = g.W / r.W;
    const double normal_r = r.normal;
    const double normal_power = alpha * normal_r;
    const double shot_noise = (alpha + 1.0) * r.S / normal_r;
    const double Wr = r.W;

    void *mesh;
    if ((s = scratch(ctx, "m.mesh", sizeof(float) * n3, &mesh)) != CM_OK) return s;
    cm_mas_desc md = o->mas;
    md.algo = CM_MAS_AUTO;
    const float *wg, *wr;
    if ((s = scaled_weights(ctx, "fkp.wg", rg, n, 1.0, &wg)) != CM_OK) return s;
    if ((s = mesh_assign_impl(ctx, &md, pos, wg, n, mesh, CM_F32, 0)) != CM_OK) return s;
    if ((s = scaled_weights(ctx, "fkp.wr", rr, nr, -alpha, &wr)) != CM_OK) return s;
    if ((s = assign_randoms(ctx, &md, rpos, wr, nr, mesh, 1)) != CM_OK) return s;
    std::vector<double> hs;
    std::vector<unsigned long long> hc;
    if ((s = shells_of_mesh(ctx, o, mesh, 0.0, hs, hc)) != CM_OK) return s;
    const double sn = o->sn_correction ? shot_noise : 0.0;
    pk_finalize(o->nshells, normal_power, sn * normal_power, hs.data(), hc.data(), pk0, pk2, pk4, nmodes);
    if (info) {
        info->n_obj = g.W;
        info->shot_noise = shot_noise;
        info->normal_power = normal_power;
        info->tweb = cm_tweb_stats{};
    }
    if (f->window && pk_window) {
        // the window's mode set (FftwFunctions.cpp:1248-1262, 1354-1568): the
        // stored modes outside the first octant, plus the zero mode
        // |W(0)|^2 = (mesh sum)^2 in the shell of k = 0 (shell 0 when kmin = 0
        // under linear binning, always under log binning, :1110-1113).  The
        // mean is removed on load (float32 transform of a small residual) and
        // the zero mode added here from the FP64 mesh sum.
        const float *wr1;
        if ((s = scaled_weights(ctx, "fkp.wr", rr, nr, 1.0, &wr1)) != CM_OK) return s;
        if ((s = assign_randoms(ctx, &md, rpos, wr1, nr, mesh, 0)) != CM_OK) return s;
        double msum = 0.0;
        if ((s = shells_of_mesh(ctx, o, mesh, Wr / (double)n3, hs, hc, true, &msum)) != CM_OK) return s;
        if (o->log_binning || o->kmin == 0.0) {
            hs[0] += msum * msum;
            hc[0] += 1;
        }
        std::vector<double> q2((size_t)o->nshells), q4((size_t)o->nshells);
        std::vector<int64_t> nm((size_t)o->nshells);
        pk_finalize(o->nshells, normal_power / (alpha * alpha), 0.0, hs.data(), hc.data(), pk_window, q2.data(),
                    q4.data(), nm.data());
    }
    return CM_OK;
}

}  // namespace cm

using namespace cm;

extern "C" cm_status cm_measure_fkp(cm_ctx *ctx, const cm_measure_opts *o, const cm_fkp_opts *f, const void *pos,
                                    const void *w, size_t n, const void *rpos, const void *rw, size_t nr, double *pk0,
                                    double *pk2, double *pk4, int64_t *nmodes, double *pk_window,
                                    cm_measure_info *info) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !f || (!pos && n) || (!rpos && nr)) return fail(ctx, CM_ERR_INVALID, "cm_measure_fkp: null pointer");
    if (!fft_size_ok(o->mas.N)) return fail(ctx, CM_ERR_INVALID, "cm_measure_fkp: N=%d not supported", o->mas.N);
    if (o->mas.pos_dtype != CM_F32) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_measure_fkp: float32 positions only");
    cm_status sb = check_binning(ctx, o->nshells, o->kmin, o->dk, o->log_binning, o->dlog, o->float_shell_rule,
                                 "cm_measure_fkp");
    if (sb != CM_OK) return sb;
    if (o->do_tweb || o->interlace)
        return fail(ctx, CM_ERR_UNSUPPORTED, "cm_measure_fkp: the T-web and interlacing options must be off");
    TuneScope ts(ctx, o->mas.N);
    return measure_fkp_impl(ctx, o, f, pos, (const float *)w, n, rpos, (const float *)rw, nr, pk0, pk2, pk4, nmodes,
                            pk_window, info);
}
