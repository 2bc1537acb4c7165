// Host implementation of the reference-named interface (see cosmicatlas_host.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "cosmicatlas_host.h"

namespace cmh {

static cm_ctx *g_ctx = nullptr;

cm_ctx *ctx() {
    if (!g_ctx) {
        cm_status s = cm_ctx_create(0, 0, &g_ctx);
        if (s != CM_OK) throw Error(std::string("cannot open the GPU context: ") + cm_status_string(s));
    }
    return g_ctx;
}

void check(cm_status s, const char *what) {
    if (s != CM_OK) throw Error(std::string(what) + ": " + cm_last_error(g_ctx));
}

// Device buffer owned for the duration of one call.
struct DevBuf {
    void *p = nullptr;
    DevBuf(size_t bytes) { check(cm_dev_alloc(ctx(), bytes, &p), "cm_dev_alloc"); }
    ~DevBuf() {
        if (p) cm_dev_free(ctx(), p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

static void h2d(void *d, const void *h, size_t b) { check(cm_copy_h2d(ctx(), d, h, b), "cm_copy_h2d"); }
static void d2h(void *h, const void *d, size_t b) { check(cm_copy_d2h(ctx(), h, d, b), "cm_copy_d2h"); }

}  // namespace cmh

using cmh::DevBuf;
using cmh::check;
using cmh::ctx;

// ======================================================================= Params
Params::Params() { init_pars(); }

Params::Params(const std::string &par_file) {
    init_pars();
    read_pars(par_file);
    derived_pars();
}

void Params::init_pars() { raw.clear(); }

static bool to_bool(const std::string &v) { return v == "true"; }

void Params::read_pars(const std::string &file) {
    std::ifstream fin(file.c_str());
    if (fin.fail()) throw cmh::Error("Error at opening the parameter file " + file);
    std::string line;
    while (std::getline(fin, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls(line);
        std::string name, eq, value;
        ls >> name;
        if (name.empty()) continue;
        ls >> eq;
        if (eq != "=")
            throw cmh::Error("Error in parameters file at " + name + " = ???\nSymbol '=' not properly written");
        ls >> value;
        if (value.empty()) {
            std::cerr << "Warning from parameter file: value of " << name << " not specified in " << file
                      << "; assuming a default value" << std::endl;
            continue;
        }
        raw[name] = value;
        if (name == "Nft") Nft = std::stoul(value);
        else if (name == "Lbox") Lbox = (real_prec)std::atof(value.c_str());
        else if (name == "mass_assignment_scheme") {
            if (value != "NGP" && value != "CIC" && value != "TSC" && value != "PCS")
                throw cmh::Error("Unknown mass assigment scheme");
            mass_assignment_scheme = value;
        } else if (name == "MAS_correction") MAS_correction = to_bool(value);
        else if (name == "type_of_binning") {
            if (value != "linear" && value != "log") throw cmh::Error("Unknown type of binning");
            type_of_binning = value;
        } else if (name == "DeltaKmin") DeltaKmin = (real_prec)std::atof(value.c_str());
        else if (name == "ndel_data") ndel_data = std::atoi(value.c_str());
        else if (name == "N_log_bins") N_log_bins = std::atoi(value.c_str());
        else if (name == "SN_correction") SN_correction = to_bool(value);
        else if (name == "lambdath") lambdath = (real_prec)std::atof(value.c_str());
        else if (name == "smscale") smscale = (real_prec)std::atof(value.c_str());
        else if (name == "Statistics") statistics = value;
        else if (name == "Input_type") input_type = value;
        else if (name == "Output_directory") Output_directory = value;
        else if (name == "file_power") file_power = value;
        else if (name == "Name_survey") Name_survey = value;
        else if (name == "file_bispectrum") file_bispectrum = value;
        else if (name == "kmax_bk") kmax_bk = std::atof(value.c_str());
        else if (name == "kmin_bk") kmin_bk = std::atof(value.c_str());
        else if (name == "delta_grid_file") delta_grid_file = value;
        else if (name == "Catalogue_file") file_catalogue = value;
        else if (name == "Input_dir_cat") Input_dir_cat = value;
        else if (name == "use_random_catalog") use_random_catalog = to_bool(value);
        else if (name == "Random_file") file_random = value;
        else if (name == "file_window") file_window = value;
        else if (name == "i_coord1_r") i_coord1_r = std::atoi(value.c_str());
        else if (name == "i_coord2_r") i_coord2_r = std::atoi(value.c_str());
        else if (name == "i_coord3_r") i_coord3_r = std::atoi(value.c_str());
        else if (name == "i_weight1_r") i_weight1_r = std::atoi(value.c_str());
        else if (name == "i_mean_density_g") i_mean_density_g = std::atoi(value.c_str());
        else if (name == "i_mean_density_r") i_mean_density_r = std::atoi(value.c_str());
        else if (name == "FKP_weight") FKP_weight = to_bool(value);
        else if (name == "Pest") Pest = (real_prec)std::atoi(value.c_str());  // read as an integer (Params.cpp:3026)
        else if (name == "measure_cross") measure_cross = to_bool(value);
        else if (name == "i_coord1_g") i_coord1_g = std::atoi(value.c_str());
        else if (name == "i_coord2_g") i_coord2_g = std::atoi(value.c_str());
        else if (name == "i_coord3_g") i_coord3_g = std::atoi(value.c_str());
        else if (name == "i_weight1_g") i_weight1_g = std::atoi(value.c_str());
        else if (name == "i_v1_g") i_v1_g = std::atoi(value.c_str());
        else if (name == "i_v2_g") i_v2_g = std::atoi(value.c_str());
        else if (name == "i_v3_g") i_v3_g = std::atoi(value.c_str());
        else if (name == "vel_units_g") vel_units_g = value;
        else if (name == "redshift_space_coords_g") redshift_space_coords_g = to_bool(value);
        else if (name == "Redshift") redshift = std::atof(value.c_str());
        else if (name == "file_bin_x_coord") file_bin_x_coord = value;
        else if (name == "file_bin_y_coord") file_bin_y_coord = value;
        else if (name == "file_bin_z_coord") file_bin_z_coord = value;
        else if (name == "N_lines_binary") N_lines_binary = std::strtoul(value.c_str(), nullptr, 10);
        else if (name == "sys_of_coord_g") sys_of_coord_g = std::atoi(value.c_str());
        else if (name == "angles_units_g") angles_units_g = value;
        else if (name == "om_matter") om_matter = std::atof(value.c_str());
        else if (name == "om_radiation") om_radiation = std::atof(value.c_str());
        else if (name == "om_vac") om_vac = std::atof(value.c_str());
        else if (name == "om_k") om_k = std::atof(value.c_str());
        else if (name == "wde_eos") wde_eos = std::atof(value.c_str());
        else if (name == "Hubble") Hubble = std::atof(value.c_str());
        else if (name == "Nbins_redshift") Nbins_redshift = std::atoi(value.c_str());
        else if (name == "redshift_min_sample") redshift_min_sample = std::atof(value.c_str());
        else if (name == "redshift_max_sample") redshift_max_sample = std::atof(value.c_str());
        else if (name == "shell_rule") {
            if (value != "exact" && value != "float") throw cmh::Error("shell_rule must be exact or float");
            shell_rule = value;
        } else if (name == "compute_tweb") compute_tweb = to_bool(value);
        // unknown keys are ignored, like the reference
    }
}

void Params::derived_pars() {
    // Params.cpp:3531-3542
    mass_assignment = mass_assignment_scheme == "NGP" ? 0 : mass_assignment_scheme == "CIC" ? 1
                      : mass_assignment_scheme == "TSC" ? 2 : 3;
    // Params.cpp:4079-4115
    NGRID = Nft * Nft * Nft;
    NGRID_h = Nft * Nft * (Nft / 2 + 1);
    delta_x = (double)Lbox / (double)Nft;
    deltak_x = 2.0 * M_PI / (double)Lbox;
    deltak_0 = std::sqrt(3.0 * deltak_x * deltak_x) / std::sqrt(3.0);
    kmin = DeltaKmin * deltak_0;
    kmax = std::sqrt(3.0 * std::pow(0.5 * Nft * deltak_x, 2)) / std::sqrt(3.0);
    DeltaK_data = ndel_data * deltak_0;
    Nnp_data = type_of_binning == "log" ? N_log_bins : (int)(Nft / ndel_data / 2);
    Deltal = std::log10(kmax / (kmin > 0 ? kmin : deltak_0)) / (double)N_log_bins;
    Xoffset = Yoffset = Zoffset = Lbox / 2;  // Params.cpp:3612
}

std::string Params::get(const std::string &key, const std::string &def) const {
    auto it = raw.find(key);
    return it == raw.end() ? def : it->second;
}

void Params::show_params() const {
    std::printf("Nft = %lu\nLbox = %.9g\nNGRID = %lu\nNGRID_h = %lu\nmass_assignment_scheme = %s\n"
                "mass_assignment = %d\nMAS_correction = %s\ntype_of_binning = %s\nDeltaKmin = %.9g\n"
                "ndel_data = %d\nSN_correction = %s\nlambdath = %.9g\nsmscale = %.9g\nStatistics = %s\n"
                "Input_type = %s\nOutput_directory = %s\nfile_power = %s\nName_survey = %s\n"
                "delta_x = %.17g\ndeltak_x = %.17g\ndeltak_0 = %.17g\nkmin = %.17g\nkmax = %.17g\n"
                "DeltaK_data = %.17g\nNnp_data = %d\nshell_rule = %s\ncompute_tweb = %s\n",
                Nft, (double)Lbox, NGRID, NGRID_h, mass_assignment_scheme.c_str(), mass_assignment,
                MAS_correction ? "true" : "false", type_of_binning.c_str(), (double)DeltaKmin, ndel_data,
                SN_correction ? "true" : "false", (double)lambdath, (double)smscale, statistics.c_str(),
                input_type.c_str(), Output_directory.c_str(), file_power.c_str(), Name_survey.c_str(), delta_x,
                deltak_x, deltak_0, kmin, kmax, DeltaK_data, Nnp_data, shell_rule.c_str(),
                compute_tweb ? "true" : "false");
    if (sys_of_coord_g == 1) {
        std::vector<double> zc, rc;
        redshift_table(*this, zc, rc);
        std::printf("zcosmo =");
        for (double v : zc) std::printf(" %.17g", v);
        std::printf("\nrcosmo =");
        for (double v : rc) std::printf(" %.17g", v);
        std::printf("\n");
    }
    if (i_v3_g > 0)
        std::printf("redshift_space_coords_g = %s\nrsd_conversion_factor = %.9g\n",
                    redshift_space_coords_g ? "true" : "false", (double)(real_prec)rsd_conversion_factor(*this));
}

// =============================================================== mass assignment
static void density_soa(int scheme, ULONG N, real_prec L, real_prec min1, real_prec min2, real_prec min3,
                        const std::vector<real_prec> &xp, const std::vector<real_prec> &yp,
                        const std::vector<real_prec> &zp, const std::vector<real_prec> &mass,
                        std::vector<real_prec> &delta, bool weightmass) {
    const size_t n = xp.size();
    if (yp.size() != n || zp.size() != n) throw cmh::Error("getDensity: coordinate arrays differ in size");
    if (weightmass && mass.size() != n) throw cmh::Error("getDensity: mass array size");
    const size_t n3 = (size_t)N * N * N;
    std::vector<real_prec> soa(3 * n);
    std::copy(xp.begin(), xp.end(), soa.begin());
    std::copy(yp.begin(), yp.end(), soa.begin() + n);
    std::copy(zp.begin(), zp.end(), soa.begin() + 2 * n);
    DevBuf dpos(sizeof(real_prec) * soa.size() + 4), dmesh(sizeof(real_prec) * n3);
    cmh::h2d(dpos.p, soa.data(), sizeof(real_prec) * soa.size());
    void *dw = nullptr;
    DevBuf dmass(weightmass ? sizeof(real_prec) * n + 4 : 4);
    if (weightmass) {
        cmh::h2d(dmass.p, mass.data(), sizeof(real_prec) * n);
        dw = dmass.p;
    }
    cm_mas_desc d;
    d.N = (int)N;
    d.L = L;
    d.xmin[0] = min1;
    d.xmin[1] = min2;
    d.xmin[2] = min3;
    d.layout = CM_SOA;
    d.pos_dtype = CM_F32;
    d.scheme = (cm_mas)scheme;
    d.interlace_shift = 0.0;
    d.algo = CM_MAS_AUTO;
    // the reference adds each particle's stencil onto delta (DELTAc(i,j,k) +=,
    // Miscelanious.cpp:454-557): a mesh of the right size is accumulated onto,
    // any other is replaced by the deposit alone
    const bool acc = delta.size() == n3;
    if (acc) cmh::h2d(dmesh.p, delta.data(), sizeof(real_prec) * n3);
    check(cm_mesh_assign(ctx(), &d, dpos.p, dw, n, dmesh.p, CM_F32, acc ? 1 : 0), "cm_mesh_assign");
    delta.resize(n3);
    cmh::d2h(delta.data(), dmesh.p, sizeof(real_prec) * n3);
}

#define GETDENSITY(NAME, SCHEME)                                                                                \
    void NAME(ULONG N1, ULONG N2, ULONG N3, real_prec L1, real_prec L2, real_prec L3, real_prec, real_prec,     \
              real_prec, real_prec min1, real_prec min2, real_prec min3, const std::vector<real_prec> &xp,     \
              const std::vector<real_prec> &yp, const std::vector<real_prec> &zp,                              \
              const std::vector<real_prec> &Par_mass, std::vector<real_prec> &delta, bool weightmass) {        \
        if (N1 != N2 || N2 != N3 || L1 != L2 || L2 != L3) throw cmh::Error(#NAME ": cubic meshes only");       \
        density_soa(SCHEME, N1, L1, min1, min2, min3, xp, yp, zp, Par_mass, delta, weightmass);                \
    }
GETDENSITY(getDensity_NGP, 0)
GETDENSITY(getDensity_CIC, 1)
GETDENSITY(getDensity_TSC, 2)

void grid_assignment(const Params &p, const std::vector<real_prec> &xyz, const std::vector<real_prec> &w,
                     std::vector<real_prec> &field) {
    const size_t n = xyz.size() / 3;
    const size_t n3 = p._NGRID();
    DevBuf dpos(sizeof(real_prec) * xyz.size() + 4), dmesh(sizeof(real_prec) * n3);
    cmh::h2d(dpos.p, xyz.data(), sizeof(real_prec) * xyz.size());
    DevBuf dw(w.empty() ? 4 : sizeof(real_prec) * w.size());
    if (!w.empty()) cmh::h2d(dw.p, w.data(), sizeof(real_prec) * w.size());
    cm_mas_desc d;
    d.N = (int)p._Nft();
    d.L = p._Lbox();
    // x -= Xoffset - L/2 (Miscelanious.cpp:2604-2606)
    d.xmin[0] = (double)p._Xoffset() - 0.5 * p._Lbox();
    d.xmin[1] = (double)p._Yoffset() - 0.5 * p._Lbox();
    d.xmin[2] = (double)p._Zoffset() - 0.5 * p._Lbox();
    d.layout = CM_AOS;
    d.pos_dtype = CM_F32;
    d.scheme = (cm_mas)p._mass_assignment();
    d.interlace_shift = 0.0;
    d.algo = CM_MAS_AUTO;
    check(cm_mesh_assign(ctx(), &d, dpos.p, w.empty() ? nullptr : dw.p, n, dmesh.p, CM_F32, 0), "cm_mesh_assign");
    field.resize(n3);
    cmh::d2h(field.data(), dmesh.p, sizeof(real_prec) * n3);
}

real_prec get_nobjects(const std::vector<real_prec> &f) {
    DevBuf d(sizeof(real_prec) * f.size() + 4);
    cmh::h2d(d.p, f.data(), sizeof(real_prec) * f.size());
    double s = 0;
    check(cm_mesh_sum(ctx(), d.p, CM_F32, f.size(), &s), "cm_mesh_sum");
    return (real_prec)s;
}

// ======================================================================== FFT
void do_fftw_r2c(ULONG Nft, std::vector<real_prec> &in, complex_prec *out) {
    const size_t n3 = (size_t)Nft * Nft * Nft, nh = (size_t)Nft * Nft * (Nft / 2 + 1);
    if (in.size() < n3) throw cmh::Error("do_fftw_r2c: input smaller than Nft^3");
    DevBuf din(sizeof(real_prec) * n3), dout(sizeof(complex_prec) * nh);
    cmh::h2d(din.p, in.data(), sizeof(real_prec) * n3);
    check(cm_fft_r2c(ctx(), (int)Nft, CM_F32, din.p, dout.p, 0.0), "cm_fft_r2c");
    cmh::d2h(out, dout.p, sizeof(complex_prec) * nh);
}

void do_fftw_c2r(int Nft, complex_prec *in, std::vector<real_prec> &out) {
    const size_t n3 = (size_t)Nft * Nft * Nft, nh = (size_t)Nft * Nft * (Nft / 2 + 1);
    DevBuf din(sizeof(complex_prec) * nh), dout(sizeof(real_prec) * n3);
    cmh::h2d(din.p, in, sizeof(complex_prec) * nh);
    check(cm_fft_c2r(ctx(), Nft, CM_F32, din.p, dout.p), "cm_fft_c2r");
    out.resize(n3);
    cmh::d2h(out.data(), dout.p, sizeof(real_prec) * n3);
}

// ======================================================================= T-web
void PoissonSolver(real_prec Lbox, ULONG Nft, std::vector<real_prec> &in, std::vector<real_prec> &out) {
    const size_t n3 = (size_t)Nft * Nft * Nft;
    DevBuf din(sizeof(real_prec) * n3), dout(sizeof(real_prec) * n3);
    cmh::h2d(din.p, in.data(), sizeof(real_prec) * n3);
    cm_tweb_opts o{(int)Nft, (double)Lbox, 0.0, 0.0, 0.0, 1.0};
    check(cm_poisson(ctx(), &o, (const float *)din.p, (float *)dout.p), "cm_poisson");
    out.resize(n3);
    cmh::d2h(out.data(), dout.p, sizeof(real_prec) * n3);
}

void EigenValuesTweb(ULONG Nft, real_prec L1, std::vector<real_prec> &, std::vector<real_prec> &phi,
                     std::vector<real_prec> &out1, std::vector<real_prec> &out2, std::vector<real_prec> &out3) {
    const size_t n3 = (size_t)Nft * Nft * Nft;
    DevBuf dphi(sizeof(real_prec) * n3), dlab(n3), dlam(3 * sizeof(real_prec) * n3);
    cmh::h2d(dphi.p, phi.data(), sizeof(real_prec) * n3);
    cm_tweb_opts o{(int)Nft, (double)L1, 0.0, 0.0, 0.0, 1.0};
    check(cm_tweb_from_phi(ctx(), &o, (const float *)dphi.p, (int8_t *)dlab.p, (float *)dlam.p, nullptr),
          "cm_tweb_from_phi");
    out1.resize(n3);
    out2.resize(n3);
    out3.resize(n3);
    cmh::d2h(out1.data(), dlam.p, sizeof(real_prec) * n3);
    cmh::d2h(out2.data(), (char *)dlam.p + sizeof(real_prec) * n3, sizeof(real_prec) * n3);
    cmh::d2h(out3.data(), (char *)dlam.p + 2 * sizeof(real_prec) * n3, sizeof(real_prec) * n3);
}

// classes of a combo code as a 4-bit mask (bit c-1 for class c)
static unsigned combo_classes(int code) {
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

// Source/Cwclass.cpp:1925-1968: code 0 (or a list {0}) takes every cell
bool Cwclass::get_cell_classified(int sua, ULONG ig) const {
    if (cwt_used.size() == 1 && cwt_used[0] == 0) return true;
    const int code = cwt_used[(size_t)sua];
    if (code == 0) return true;
    const int lab = CWClass[ig];
    return lab >= 1 && lab <= 4 && (combo_classes(code) >> (lab - 1) & 1u);
}

// Source/Cwclass.cpp:1983-2030: one code -> the code; several -> index of the
// last code containing the cell's class, else 0
int Cwclass::get_Tclassification(ULONG ig) const {
    if (cwt_used.size() == 1) return cwt_used[0];
    int ans = 0;
    const int lab = CWClass[ig];
    for (size_t ic = 0; ic < cwt_used.size(); ++ic)
        if (lab >= 1 && lab <= 4 && (combo_classes(cwt_used[ic]) >> (lab - 1) & 1u)) ans = (int)ic;
    return ans;
}

void Cwclass::get_CWC(std::vector<real_prec> &delta) {
    const ULONG N = params._Nft();
    const size_t n3 = (size_t)N * N * N;
    DevBuf dd(sizeof(real_prec) * n3), dlab(n3), dlam(3 * sizeof(real_prec) * n3);
    cmh::h2d(dd.p, delta.data(), sizeof(real_prec) * n3);
    // margin: the cells whose label the float32 rounding of the potential can
    // change are counted (tol = K(N) u ||phi||_rms / d^2)
    cm_tweb_opts o{(int)N, (double)params._Lbox(), (double)params._smscale(), (double)params._lambdath(),
                   CM_TWEB_TOL_ROUNDING, 1.0, spectral_hessian ? CM_HESSIAN_SPECTRAL : CM_HESSIAN_FD};
    cm_tweb_stats st;
    check(cm_tweb(ctx(), &o, (const float *)dd.p, (int8_t *)dlab.p, (float *)dlam.p, &st), "cm_tweb");
    std::vector<int8_t> lab(n3);
    cmh::d2h(lab.data(), dlab.p, n3);
    CWClass.assign(lab.begin(), lab.end());
    lambda1.resize(n3);
    lambda2.resize(n3);
    lambda3.resize(n3);
    cmh::d2h(lambda1.data(), dlam.p, sizeof(real_prec) * n3);
    cmh::d2h(lambda2.data(), (char *)dlam.p + sizeof(real_prec) * n3, sizeof(real_prec) * n3);
    cmh::d2h(lambda3.data(), (char *)dlam.p + 2 * sizeof(real_prec) * n3, sizeof(real_prec) * n3);
    // volume fractions (Cwclass.cpp:891-894; the reference swaps the sheet and
    // filament names at :987-988 -- reported correctly here)
    volume_knots = (real_prec)st.counts[0] / (real_prec)n3;
    volume_filaments = (real_prec)st.counts[1] / (real_prec)n3;
    volume_sheets = (real_prec)st.counts[2] / (real_prec)n3;
    volume_voids = (real_prec)st.counts[3] / (real_prec)n3;
    n_marginal = st.n_marginal;
    margin_tol = st.margin_tol;
}

// =============================================================== FftwFunctions
FftwFunctions::~FftwFunctions() { free_fftw_vectors(); }

void FftwFunctions::resize_fftw_vectors() {
    free_fftw_vectors();
    data_g.assign(params._NGRID(), 0);
    data_out_g = new complex_prec[params._NGRID_h()];
}

void FftwFunctions::free_fftw_vectors() {
    delete[] data_out_g;
    data_out_g = nullptr;
}

// FftwFunctions.cpp:159-174
void FftwFunctions::raw_sampling(real_prec vol) {
    normal_p = n_gal * n_gal / vol;
    s_r = vol / n_gal;
    alpha = 1.0;
    normal_power = normal_p;
    shot_noise = s_r;
}

// FftwFunctions.cpp:177-212: the random branch from the catalogue sums
// (alpha = W_g/W_r, normal_power = alpha normal_p, shot noise
// (alpha + 1) S_r/normal_p, window normalisation normal_power/alpha^2), the box
// branch from raw_sampling
void FftwFunctions::get_parameters_estimator(bool verbose) {
    if (params._use_random_catalog()) {
        if (!(w_r > 0) || !(normal_p > 0))
            throw cmh::Error("use_random_catalog: the random catalogue has no weight (W_r = 0 or sum nbar w^2 = 0)");
        alpha = (real_prec)((double)w_g / (double)w_r);
        normal_power = alpha * normal_p;
        shot_noise = (real_prec)((alpha + 1.0) * (double)s_r / (double)normal_p);
        normal_window = normal_power / (alpha * alpha);
        if (verbose)
            std::printf("** Ngal = %.9g\n** Nran = %.9g\n** Wg = %.9g\n** Wr = %.9g\n** Sg = %.9g\n** Sr = %.9g\n"
                        "** alpha = %.9g\n** Normalization = %.9g\n** Shot Noise = %.9g\n",
                        (double)n_gal, (double)n_ran, (double)w_g, (double)w_r, (double)s_g, (double)s_r,
                        (double)alpha, (double)normal_power, (double)shot_noise);
        return;
    }
    alpha = 1.0;
    shot_noise = s_r;
    normal_window = 1.0;
    if (verbose)
        std::printf("** Ngal = %.9g\n** Normalization = %.9g\n** Shot Noise = %.9g\n", (double)n_gal,
                    (double)normal_power, (double)shot_noise);
}

// FftwFunctions.cpp:275-294: delta = rho_g - alpha rho_r with randoms, else the
// mean density removed
void FftwFunctions::get_fluctuation() {
    if (params._use_random_catalog()) {
        if (data_r.size() != data_g.size()) throw cmh::Error("get_fluctuation: the random mesh is missing");
        for (size_t i = 0; i < data_g.size(); ++i) data_g[i] -= alpha * data_r[i];
        return;
    }
    const real_prec m = n_gal / (real_prec)params._NGRID();
    for (auto &v : data_g) v -= m;
}

void FftwFunctions::shells(const std::vector<real_prec> &mesh, bool window, std::vector<real_prec> &p0,
                           std::vector<real_prec> *p2, std::vector<real_prec> *p4, std::vector<int> &mod) {
    const ULONG N = params._Nft();
    const size_t n3 = (size_t)N * N * N, nh = params._NGRID_h();
    DevBuf din(sizeof(real_prec) * n3), dk(sizeof(complex_prec) * nh);
    cmh::h2d(din.p, mesh.data(), sizeof(real_prec) * n3);
    check(cm_fft_r2c(ctx(), (int)N, CM_F32, din.p, dk.p, 0.0), "cm_fft_r2c");
    if (data_out_g && !window) cmh::d2h(data_out_g, dk.p, sizeof(complex_prec) * nh);
    const int ns = params._d_Nnp_data();
    const bool logb = params._type_of_binning() == "log";
    // the log rule divides by kmin (FftwFunctions.cpp:1003, rkmin = 1/kmin)
    if (logb && !(params._d_kmin() > 0)) throw cmh::Error("type_of_binning = log needs DeltaKmin > 0");
    cm_pk_opts o;
    o.N = (int)N;
    o.L = params._Lbox();
    o.nshells = ns;
    o.kmin = params._d_kmin();
    o.dk = params._d_DeltaK_data();
    o.scheme = (cm_mas)params._mass_assignment();
    o.mas_correction = params._MAS_correction();
    o.shot_noise = (params._SN_correction() && !window) ? shot_noise : 0.0;
    o.normal_power = window ? normal_window : normal_power;
    o.float_shell_rule = logb ? 0 : params._float_shell_rule();
    o.log_binning = logb ? 1 : 0;
    o.dlog = params._d_Deltal();
    o.window_set = window ? 1 : 0;
    std::vector<double> a0(ns), a2(ns), a4(ns);
    std::vector<int64_t> nm(ns);
    check(cm_pk_shells(ctx(), &o, CM_F32, dk.p, a0.data(), a2.data(), a4.data(), nm.data()), "cm_pk_shells");
    p0.assign(a0.begin(), a0.end());
    if (p2) p2->assign(a2.begin(), a2.end());
    if (p4) p4->assign(a4.begin(), a4.end());
    mod.assign(nm.begin(), nm.end());
}

void FftwFunctions::get_power_spectrum_fkp(std::vector<real_prec> &power_g0, std::vector<real_prec> &power_g2,
                                           std::vector<real_prec> &power_g4, std::vector<real_prec> &power_r,
                                           std::vector<int> &mod_g, std::vector<int> &mod_r) {
    shells(data_g, false, power_g0, &power_g2, &power_g4, mod_g);
    if (params._use_random_catalog()) shells(data_r, true, power_r, nullptr, nullptr, mod_r);
}

void FftwFunctions::get_power_spectrum_fkp(std::vector<real_prec> &power_g0, std::vector<real_prec> &power_g2,
                                           std::vector<real_prec> &power_g4, std::vector<int> &mod_g) {
    shells(data_g, false, power_g0, &power_g2, &power_g4, mod_g);
}

void FftwFunctions::power_spectrum(std::vector<real_prec> &powerk_g0, std::vector<int> &mod_g) {
    shells(data_g, false, powerk_g0, nullptr, nullptr, mod_g);
}

void FftwFunctions::get_bispectrum_fkp(char, std::vector<real_prec> &bispect, std::vector<real_prec> &sn_bispect,
                                       std::vector<int> &mod) {
    const ULONG N = params._Nft();
    const size_t n3 = (size_t)N * N * N, nh = params._NGRID_h();
    if (params._type_of_binning() != "linear")
        throw cmh::Error("Bk_fkp: log binning of the bispectrum is not part of this build");
    const int nb = params._d_Nnp_data();
    if (nb > 64) throw cmh::Error("Bk_fkp: at most 64 k bins (Nft/(2 ndel_data) = " + std::to_string(nb) + ")");
    DevBuf din(sizeof(real_prec) * n3), dk(sizeof(complex_prec) * nh);
    cmh::h2d(din.p, data_g.data(), sizeof(real_prec) * n3);
    check(cm_fft_r2c(ctx(), (int)N, CM_F32, din.p, dk.p, 0.0), "cm_fft_r2c");
    cm_bkd_opts o;
    o.N = (int)N;
    o.L = params._Lbox();
    o.nbins = nb;
    o.kmin = params._d_kmin();
    o.dk = params._d_DeltaK_data();
    o.scheme = (cm_mas)params._mass_assignment();
    o.mas_correction = params._MAS_correction();
    o.sn_correction = params._SN_correction();
    o.n_obj = n_gal;
    const size_t nb3 = (size_t)nb * nb * nb;
    std::vector<double> bk(nb3), sn(nb3), snsum(nb3), pk(nb);
    std::vector<int64_t> nt(nb3), nk(nb);
    check(cm_bispectrum_direct(ctx(), &o, dk.p, bk.data(), sn.data(), nt.data(), pk.data(), nk.data(), nullptr,
                               snsum.data()),
          "cm_bispectrum_direct");
    bispect.assign(n3, 0);
    sn_bispect.assign(n3, 0);
    mod.assign(n3, 0);
    for (int a = 0; a < nb; ++a)
        for (int b = a; b < nb; ++b)
            for (int c = b; c < nb; ++c) {
                const size_t e = ((size_t)a * nb + b) * nb + c, I = ((size_t)a * N + b) * N + c;
                bispect[I] = (real_prec)bk[e];
                sn_bispect[I] = (real_prec)snsum[e];
                mod[I] = (int)nt[e];
            }
}

void FftwFunctions::get_bispectrum_fkp_fast(std::vector<real_prec> &pk, std::vector<real_prec> &bispect,
                                            std::vector<int> &mod, const std::string &file) {
    const ULONG N = params._Nft();
    const size_t n3 = (size_t)N * N * N, nh = params._NGRID_h();
    if (params._type_of_binning() != "linear")
        throw cmh::Error("Bk_fkp_fast: log binning of the bispectrum is not part of this build");
    // FftwFunctions::set_params (:5064-5095): DeltaK_Bis = DeltaK_data, Nshells_bk =
    // int(kmax_bk / DeltaK_Bis), kmin_bk raised to the fundamental mode; the
    // triangle loop starts at floor((kmin_bk - kmin)/DeltaK_Bis) (:3524)
    const double dkb = params._d_DeltaK_data(), kf = 2.0 * M_PI / params._Lbox();
    const int ns = (int)(params._kmax_bk() / dkb);
    const double kmin_bk = std::max(params._kmin_bk(), kf);
    const int smin = std::max(0, (int)std::floor((kmin_bk - params._d_kmin()) / dkb));
    if (ns < 1 || smin >= ns) throw cmh::Error("Bk_fkp_fast: no shells between kmin_bk and kmax_bk");
    DevBuf din(sizeof(real_prec) * n3), dk(sizeof(complex_prec) * nh);
    cmh::h2d(din.p, data_g.data(), sizeof(real_prec) * n3);
    check(cm_fft_r2c(ctx(), (int)N, CM_F32, din.p, dk.p, 0.0), "cm_fft_r2c");
    cm_bk_opts o;
    o.N = (int)N;
    o.L = params._Lbox();
    o.nshells = ns;
    o.kmin = params._d_kmin();
    o.dk = dkb;
    o.shell_min = smin;
    o.scheme = (cm_mas)params._mass_assignment();
    o.mas_correction = params._MAS_correction();
    o.sn_correction = params._SN_correction();
    o.n_obj = n_gal;
    const int cap = 8 * 256;
    std::vector<int32_t> ijk(3 * (size_t)cap);
    std::vector<double> bk(cap), sn(cap), nt(cap);
    int nb = 0;
    check(cm_bispectrum(ctx(), &o, dk.p, cap, ijk.data(), bk.data(), sn.data(), nt.data(), &nb), "cm_bispectrum");
    std::ofstream out(file.c_str());
    if (!out) throw cmh::Error("cannot write " + file);
    out.precision(10);
    out.setf(std::ios::showpoint);
    out.setf(std::ios::scientific);
    bispect.assign((size_t)ns * ns * ns, 0);
    mod.assign((size_t)ns * ns * ns, 0);
    for (int b = 0; b < nb; ++b) {
        const int i = ijk[3 * b], j = ijk[3 * b + 1], k = ijk[3 * b + 2];
        const size_t I = ((size_t)i * ns + j) * ns + k;
        bispect[I] = (real_prec)bk[b];
        mod[I] = (int)nt[b];
        out << i << "\t" << j << "\t" << k << "\t" << bk[b] << "\t" << sn[b] << "\t" << (long long)nt[b] << "\n";
    }
    pk.assign(params._d_Nnp_data(), 0);
}

// ================================================================== file I/O
void read_array(const std::string &fname, std::vector<real_prec> &out) {
    std::ifstream f(fname.c_str(), std::ios::binary);
    if (!f) throw cmh::Error("File " + fname + " not found");
    // sized from the mesh, not read to EOF (the reference's read_array_fast
    // pushes one extra element, FileOutput.cpp:756-778)
    f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(sizeof(real_prec) * out.size()));
    if ((size_t)f.gcount() != sizeof(real_prec) * out.size())
        throw cmh::Error("File " + fname + " holds fewer than Nft^3 float32 values");
}

void write_array(const std::string &fname, const std::vector<real_prec> &in) {
    std::ofstream f((fname + ".dat").c_str(), std::ios::binary);
    if (!f) throw cmh::Error("cannot write " + fname + ".dat");
    f.write(reinterpret_cast<const char *>(in.data()), (std::streamsize)(sizeof(real_prec) * in.size()));
}

void write_to_file2(const std::string &fname, const std::vector<real_prec> &kve, const std::vector<real_prec> &pk,
                    const std::vector<int> &si, bool cero) {
    std::ofstream out(fname.c_str());
    if (!out) throw cmh::Error("cannot write " + fname);
    size_t first = (cero && !pk.empty() && pk[0] < 1) ? 1 : 0;
    for (size_t i = first; i < kve.size(); ++i) out << kve[i] << "\t" << pk[i] << "\t" << si[i] << "\n";
}

// equatorial_to_cartesian (Source/CoordinateSystem.cpp:12-18) with the angle
// units of Catalog::ang_to_cart_coordinates (Source/Catalog.cpp:13640-13652)
static void equatorial_to_cartesian(double ra, double dec, double r, double &x, double &y, double &z) {
    const double fac = M_PI / 180.0;
    x = r * std::cos(ra * fac) * std::sin(0.5 * M_PI - dec * fac);
    y = r * std::sin(ra * fac) * std::sin(0.5 * M_PI - dec * fac);
    z = r * std::cos(0.5 * M_PI - dec * fac);
}

// Hubble_function (Source/CosmologicalFunctions.cpp:47-49)
static double hubble_function(const Params &p, double z) {
    const double a = 1.0 + z;
    return p._Hubble() * std::sqrt(p._om_matter() * a * a * a + p._om_radiation() * a * a * a * a +
                                   p._om_k() * a * a + p._om_vac() * std::pow(a, 3.0 * (1.0 + p._wde_eos())));
}

// adaptive Simpson on 1/H with Richardson's correction; 1/H is smooth and
// positive on [0, z], so a relative tolerance of 1e-13 is met in a few hundred
// evaluations
static double simpson_hinv(const Params &p, double a, double b, double fa, double fm, double fb, double whole,
                           double tol, int depth) {
    const double m = 0.5 * (a + b);
    const double lm = 0.5 * (a + m), rm = 0.5 * (m + b);
    const double flm = 1.0 / hubble_function(p, lm), frm = 1.0 / hubble_function(p, rm);
    const double left = (m - a) / 6.0 * (fa + 4.0 * flm + fm);
    const double right = (b - m) / 6.0 * (fm + 4.0 * frm + fb);
    const double delta = left + right - whole;
    if (depth <= 0 || std::fabs(delta) <= 15.0 * tol) return left + right + delta / 15.0;
    return simpson_hinv(p, a, m, fa, flm, fm, left, 0.5 * tol, depth - 1) +
           simpson_hinv(p, m, b, fm, frm, fb, right, 0.5 * tol, depth - 1);
}

double comoving_distance(const Params &p, double z) {
    if (z <= 0) return 0.0;
    const double c_kms = 299792.458;  // Constants::speed_light (Headers/Constants.h:21)
    const double fa = 1.0 / hubble_function(p, 0.0), fb = 1.0 / hubble_function(p, z);
    const double fm = 1.0 / hubble_function(p, 0.5 * z);
    const double whole = z / 6.0 * (fa + 4.0 * fm + fb);
    return c_kms * simpson_hinv(p, 0.0, z, fa, fm, fb, whole, 1e-13 * z * fa, 40);
}

void redshift_table(const Params &p, std::vector<double> &zc, std::vector<double> &rc) {
    const int nz = p._Nbins_redshift();
    if (nz < 2) throw cmh::Error("Nbins_redshift must be at least 2 for the distance-redshift table");
    if (!(p._redshift_max_sample() > 0)) throw cmh::Error("redshift_max_sample must be positive");
    zc.resize(nz);
    rc.resize(nz);
    for (int i = 0; i < nz; ++i) {
        zc[i] = i * p._redshift_max_sample() / (double)(nz - 1);
        rc[i] = comoving_distance(p, zc[i]);
    }
}

// linear interpolation in the r(z) table (gsl_interp_linear, Source/Catalog.cpp:13662-13669);
// redshifts below the first node are raised to it (:13802), those above the last
// one are an error (GSL's interpolation refuses them)
static double distance_of_redshift(const std::vector<double> &zc, const std::vector<double> &rc, double z) {
    if (z < zc[0]) z = zc[0];
    if (z > zc.back())
        throw cmh::Error("redshift " + std::to_string(z) + " above the r(z) table (redshift_max_sample = " +
                         std::to_string(zc.back()) + ")");
    size_t i = (size_t)((z - zc[0]) / (zc[1] - zc[0]));
    if (i >= zc.size() - 1) i = zc.size() - 2;
    while (i > 0 && z < zc[i]) --i;
    while (i + 2 < zc.size() && z >= zc[i + 1]) ++i;
    const double t = (z - zc[i]) / (zc[i + 1] - zc[i]);
    return rc[i] + t * (rc[i + 1] - rc[i]);
}

static void to_cartesian(const Params &p, std::vector<real_prec> &xyz) {
    switch (p._sys_of_coord_g()) {
    case 0: return;
    case 1: {
        // Catalog::ang_to_cart_coordinates, I_EQZ branch (Source/Catalog.cpp:13786-13820)
        std::vector<double> zc, rc;
        redshift_table(p, zc, rc);
        const double fac = p._angles_units_g() == "R" ? 180.0 / M_PI : 1.0;
        for (size_t i = 0; i + 2 < xyz.size(); i += 3) {
            double x, y, z;
            const double r = distance_of_redshift(zc, rc, xyz[i + 2]);
            equatorial_to_cartesian(xyz[i] * fac, xyz[i + 1] * fac, r, x, y, z);
            xyz[i] = (real_prec)x;
            xyz[i + 1] = (real_prec)y;
            xyz[i + 2] = (real_prec)z;
        }
        return;
    }
    case 2: {
        const double fac = p._angles_units_g() == "R" ? 180.0 / M_PI : 1.0;  // radians -> degrees
        for (size_t i = 0; i + 2 < xyz.size(); i += 3) {
            double x, y, z;
            equatorial_to_cartesian(xyz[i] * fac, xyz[i + 1] * fac, xyz[i + 2], x, y, z);
            xyz[i] = (real_prec)x;
            xyz[i + 1] = (real_prec)y;
            xyz[i + 2] = (real_prec)z;
        }
        return;
    }
    }
    throw cmh::Error("unknown sys_of_coord_g = " + std::to_string(p._sys_of_coord_g()));
}

// three headerless real_prec files x, y, z (N_lines_binary values each, or the
// whole file when 0)
static void read_catalog_bin(const Params &p, std::vector<real_prec> &xyz) {
    std::vector<real_prec> c[3];
    const std::string f[3] = {p._file_bin_x_coord(), p._file_bin_y_coord(), p._file_bin_z_coord()};
    for (int a = 0; a < 3; ++a) {
        std::ifstream in(f[a].c_str(), std::ios::binary | std::ios::ate);
        if (!in) throw cmh::Error("File " + f[a] + " not found");
        const size_t bytes = (size_t)in.tellg();
        size_t n = bytes / sizeof(real_prec);
        if (p._N_lines_binary() > 0) {
            if (p._N_lines_binary() > n) throw cmh::Error("File " + f[a] + " holds fewer than N_lines_binary values");
            n = p._N_lines_binary();
        }
        in.seekg(0);
        c[a].resize(n);
        in.read(reinterpret_cast<char *>(c[a].data()), (std::streamsize)(n * sizeof(real_prec)));
    }
    if (c[1].size() != c[0].size() || c[2].size() != c[0].size())
        throw cmh::Error("binary coordinate files of different lengths");
    xyz.resize(3 * c[0].size());
    for (size_t i = 0; i < c[0].size(); ++i) {
        xyz[3 * i] = c[0][i];
        xyz[3 * i + 1] = c[1][i];
        xyz[3 * i + 2] = c[2][i];
    }
}

// velocity units -> comoving Mpc/h along the line of sight
// (Catalog.cpp:14088-14099, PowerSpectrumF.cpp:445-454): km/s times (1+z)/H(z);
// "alpt" velocities divided by H(z); Mpc/h as they are; any other name leaves
// the factor at 1. The velocity-bias power the reference multiplies in is 1.
double rsd_conversion_factor(const Params &p) {
    const double z = p._redshift();
    if (p._vel_units_g() == "kmps") return (1.0 + z) / hubble_function(p, z);
    if (p._vel_units_g() == "alpt") return 1.0 / hubble_function(p, z);
    return 1.0;
}

// Catalog::get_interpolated_density_field (Source/Catalog.cpp:14071-14129):
// with the coordinates in real space (redshift_space_coords_g false), every
// object moves along the line of sight, the third Cartesian axis, by its
// velocity times the conversion factor; the mesh assignment then wraps the
// moved coordinate into the box (Miscelanious.cpp:2711-2716)
static void to_redshift_space(const Params &p, std::vector<real_prec> &xyz, const std::vector<real_prec> &vlos) {
    if (p._redshift_space_coords_g()) return;
    if (vlos.size() * 3 != xyz.size()) throw cmh::Error("velocity column of a different length than the coordinates");
    const real_prec f = (real_prec)rsd_conversion_factor(p);
    for (size_t i = 0; i < vlos.size(); ++i) xyz[3 * i + 2] = xyz[3 * i + 2] + f * vlos[i];
}

void read_catalog(const Params &p, std::vector<real_prec> &xyz, std::vector<real_prec> &w) {
    xyz.clear();
    w.clear();
    if (!p._file_bin_x_coord().empty()) {
        read_catalog_bin(p, xyz);
        to_cartesian(p, xyz);
        return;
    }
    std::ifstream f(p._file_catalogue().c_str());
    if (!f) throw cmh::Error("File " + p._file_catalogue() + " not found");
    const int c[3] = {p._i_coord1_g(), p._i_coord2_g(), p._i_coord3_g()};
    const int cw = p._i_weight1_g();
    // the line-of-sight velocity: the column is read when its index is
    // positive (Catalog.cpp:2306, `if(i_vz>0)`)
    const int cv = p._i_v3_g() > 0 ? p._i_v3_g() : -1;
    int need = std::max(std::max(std::max(c[0], c[1]), std::max(c[2], cw)), cv) + 1;
    std::string line;
    std::vector<double> col;
    std::vector<real_prec> vlos;
    xyz.clear();
    w.clear();
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls(line);
        col.clear();
        double v;
        while (ls >> v) col.push_back(v);
        if ((int)col.size() < need) continue;
        for (int a = 0; a < 3; ++a) xyz.push_back((real_prec)col[c[a]]);
        if (cw >= 0) w.push_back((real_prec)col[cw]);
        if (cv >= 0) vlos.push_back((real_prec)col[cv]);
    }
    to_cartesian(p, xyz);
    if (!vlos.empty()) to_redshift_space(p, xyz, vlos);
}

void read_catalog_columns(const Params &p, const std::string &file, const int c[3], int cw, int cnb,
                          std::vector<real_prec> &xyz, std::vector<real_prec> &w, std::vector<real_prec> &nb) {
    xyz.clear();
    w.clear();
    nb.clear();
    std::ifstream f(file.c_str());
    if (!f) throw cmh::Error("File " + file + " not found");
    const int need = std::max(std::max(std::max(c[0], c[1]), std::max(c[2], cw)), cnb) + 1;
    std::string line;
    std::vector<double> col;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls(line);
        col.clear();
        double v;
        while (ls >> v) col.push_back(v);
        if ((int)col.size() < need) continue;
        for (int a = 0; a < 3; ++a) xyz.push_back((real_prec)col[c[a]]);
        if (cw >= 0) w.push_back((real_prec)col[cw]);
        if (cnb >= 0) nb.push_back((real_prec)col[cnb]);
    }
    to_cartesian(p, xyz);
}

// ============================================================== PowerSpectrumF
PowerSpectrumF::PowerSpectrumF(const Params &p) : fftw_functions(p), cwclass(p), params(p) {
    set_output_filenames();
}

// PowerSpectrumF.cpp:15-46 (real-space file name)
void PowerSpectrumF::set_output_filenames() {
    const std::string stem = params._Output_directory() + params._statistics() + "_" + params._Name_survey() +
                             "_Nft" + std::to_string(params._Nft()) + "_" + params._mass_assignment_scheme() + "_" +
                             params._file_power();
    file_power = stem + ".txt";
    file_bispectrum = params._Output_directory() + params._statistics() + "_" + params._Name_survey() + "_Nft" +
                      std::to_string(params._Nft()) + "_" + params._mass_assignment_scheme() + "_" +
                      params._file_bispectrum() + ".txt";
    file_cwc = params._Output_directory() + "CWC_" + params._Name_survey() + "_Nft" + std::to_string(params._Nft()) +
               "_" + params._mass_assignment_scheme() + "_" + params._file_power();
    file_window = params._Output_directory() + params._statistics() + "_" + params._Name_survey() + "_Nft" +
                  std::to_string(params._Nft()) + "_" + params._mass_assignment_scheme() + "_" +
                  params._file_window() + ".txt";
}

void PowerSpectrumF::kvector() {
    // PowerSpectrumF.cpp:1266-1272 / 4259-4264: bin centres kmin + dk (i+1/2);
    // log binning: kmin 10^((i-1/2) Deltal), the log centre of shell i
    // (:1273-1277 loops over the just-cleared vector and writes no k at all)
    kvector_data.resize(params._d_Nnp_data());
    const bool logb = params._type_of_binning() == "log";
    for (int i = 0; i < params._d_Nnp_data(); ++i)
        kvector_data[i] = (real_prec)(logb ? params._d_kmin() * std::pow(10.0, (i - 0.5) * params._d_Deltal())
                                           : params._d_kmin() + params._d_DeltaK_data() * (i + 0.5));
}

// Per-object sums of a catalogue (Catalog::get_interpolated_density_field,
// Source/Catalog.cpp:14142-14171): total weight w = w_user (1 without a weight
// column) times 1/(1 + Pest nbar) when FKP_weight, with nbar the object's
// mean-density column or one constant; W = sum w, S = sum w^2,
// normal = sum nbar w^2.  w receives the total weights.
struct CatalogSums {
    double n = 0, W = 0, S = 0, normal = 0;
};
static CatalogSums catalog_sums(const Params &p, size_t n, std::vector<real_prec> &w, const std::vector<real_prec> &nb,
                                double nbar_const) {
    CatalogSums s;
    std::vector<real_prec> tot(n);
    for (size_t i = 0; i < n; ++i) {
        const double nbar = nb.empty() ? nbar_const : (double)nb[i];
        double v = w.empty() ? 1.0 : (double)w[i];
        if (p._FKP_weight()) v *= 1.0 / (1.0 + (double)p._Pest() * nbar);
        tot[i] = (real_prec)v;
        s.W += v;
        s.S += v * v;
        s.normal += nbar * v * v;
    }
    s.n = (double)n;
    w.swap(tot);
    return s;
}

// -p with Input_type=catalog and use_random_catalog (PowerSpectrumF.cpp:
// 741-1363 random branch): both catalogues onto meshes with their total
// weights, alpha and the normalisations from the sums, delta = rho_g -
// alpha rho_r, P(k) with the shot noise (alpha + 1) S_r/normal_r and the
// window of the randoms.  A catalogue without a mean-density column takes the
// constant nbar = W_g / Lbox^3 (sum of the tracers' user weights over the box).
void PowerSpectrumF::compute_power_spectrum_randoms(bool verbose) {
    std::vector<real_prec> xg, wg, nbg, xr, wr, nbr;
    if (params._i_mean_density_g() >= 0) {
        const int cg[3] = {params._i_coord1_g(), params._i_coord2_g(), params._i_coord3_g()};
        read_catalog_columns(params, params._file_catalogue(), cg, params._i_weight1_g(), params._i_mean_density_g(),
                             xg, wg, nbg);
    } else {
        read_catalog(params, xg, wg);
    }
    const int cr[3] = {params._i_coord1_r(), params._i_coord2_r(), params._i_coord3_r()};
    read_catalog_columns(params, params._file_random(), cr, params._i_weight1_r(), params._i_mean_density_r(), xr, wr,
                         nbr);
    const size_t ng = xg.size() / 3, nr = xr.size() / 3;
    if (ng == 0) throw cmh::Error("the catalogue " + params._file_catalogue() + " holds no objects");
    if (nr == 0) throw cmh::Error("the random catalogue " + params._file_random() + " holds no objects");
    double wuser = 0;
    if (wg.empty()) wuser = (double)ng;
    else
        for (real_prec v : wg) wuser += v;
    const double V = (double)params._Lbox() * params._Lbox() * params._Lbox();
    const double nbar_const = wuser / V;
    const CatalogSums sg = catalog_sums(params, ng, wg, nbg, nbar_const);
    const CatalogSums sr = catalog_sums(params, nr, wr, nbr, nbar_const);
    fftw_functions.resize_fftw_vectors();
    grid_assignment(params, xg, wg, fftw_functions.data_g);
    fftw_functions.data_r.assign(params._NGRID(), 0);
    grid_assignment(params, xr, wr, fftw_functions.data_r);
    fftw_functions.set_n_gal((real_prec)sg.n);
    fftw_functions.set_w_g((real_prec)sg.W);
    fftw_functions.set_s_g((real_prec)sg.S);
    fftw_functions.set_n_ran((real_prec)sr.n);
    fftw_functions.set_w_r((real_prec)sr.W);
    fftw_functions.set_s_r((real_prec)sr.S);
    fftw_functions.set_normal_p((real_prec)sr.normal);
    fftw_functions.get_parameters_estimator(verbose);
    std::vector<real_prec> mesh;
    if (params._compute_tweb()) mesh = fftw_functions.data_g;  // kept for the T-web
    fftw_functions.get_fluctuation();
    kvector();
    if (params._statistics() != "Pk_fkp")
        throw cmh::Error("use_random_catalog: Statistics = " + params._statistics() +
                         " with randoms is not part of this build (Pk_fkp is)");
    fftw_functions.get_power_spectrum_fkp(pk0, pk2, pk4, pk_w, modes_g, modes_r);
    write_power_and_modes();
    if (params._compute_tweb()) {
        const real_prec mean = (real_prec)(sg.W / (double)params._NGRID());
        for (auto &v : mesh) v = v / mean - 1;
        cwclass.get_CWC(mesh);
        std::vector<real_prec> lab(cwclass.CWClass.begin(), cwclass.CWClass.end());
        write_array(file_cwc, lab);
    }
}

// -p with Input_type=catalog (PowerSpectrumF.cpp:741-1363; box, or the random
// branch when use_random_catalog)
void PowerSpectrumF::compute_power_spectrum(bool verbose, bool) {
    if (params._use_random_catalog()) {
        compute_power_spectrum_randoms(verbose);
        return;
    }
    std::vector<real_prec> xyz, w;
    read_catalog(params, xyz, w);
    const size_t n = xyz.size() / 3;
    if (n == 0) throw cmh::Error("the catalogue " + params._file_catalogue() + " holds no objects");
    fftw_functions.resize_fftw_vectors();
    grid_assignment(params, xyz, w, fftw_functions.data_g);
    double ngal = 0;
    if (w.empty()) ngal = (double)n;
    else
        for (real_prec v : w) ngal += v;
    fftw_functions.set_n_gal((real_prec)ngal);
    fftw_functions.raw_sampling(params._Lbox() * params._Lbox() * params._Lbox());
    fftw_functions.get_parameters_estimator(verbose);
    std::vector<real_prec> mesh = fftw_functions.data_g;  // kept for the T-web
    fftw_functions.get_fluctuation();
    kvector();
    if (params._statistics() == "Bk_fkp") {
        // PowerSpectrumF.cpp:1328-1343; the rows are the bin triples a <= b <= c
        // with triangles, "a b c k_a k_b k_c B SN_sum ntri"
        kvector_data_b.clear();
        for (int i = 0; i < params._d_Nnp_data(); ++i)
            kvector_data_b.push_back((real_prec)(params._d_kmin() + params._d_DeltaK_data() * (i + 0.5)));
        fftw_functions.get_bispectrum_fkp('d', bispectrum, sn_bispectrum, modes_tri);
        write_bispectrum();
    } else if (params._statistics() == "Bk_fkp_fast") {
        // PowerSpectrumF.cpp:1345-1356
        fftw_functions.get_bispectrum_fkp_fast(pk0, bispectrum, modes_tri, file_bispectrum);
    } else {
        fftw_functions.get_power_spectrum_fkp(pk0, pk2, pk4, modes_g);
        write_power_and_modes();
    }
    if (params._compute_tweb()) {
        const real_prec mean = (real_prec)(ngal / (double)params._NGRID());
        for (auto &v : mesh) v = v / mean - 1;
        cwclass.get_CWC(mesh);
        std::vector<real_prec> lab(cwclass.CWClass.begin(), cwclass.CWClass.end());
        write_array(file_cwc, lab);
        if (verbose)
            std::printf("T-web volume fractions: knots %.6f filaments %.6f sheets %.6f voids %.6f\n",
                        (double)cwclass.volume_knots, (double)cwclass.volume_filaments,
                        (double)cwclass.volume_sheets, (double)cwclass.volume_voids);
    }
}

// -p with Input_type=density_grid / delta_grid (PowerSpectrumF.cpp:4188-4290)
void PowerSpectrumF::compute_power_spectrum_grid() {
    std::vector<real_prec> data_in(params._NGRID(), 0);
    read_array(params._delta_grid_file(), data_in);
    fftw_functions.resize_fftw_vectors();
    const real_prec ngal = get_nobjects(data_in);
    if (params._input_type() == "density_grid") {
        const real_prec nmean = ngal / (real_prec)params._NGRID();
        for (size_t i = 0; i < data_in.size(); ++i) fftw_functions.data_g[i] = data_in[i] / nmean - 1;
    } else {
        fftw_functions.data_g = data_in;
    }
    fftw_functions.set_n_gal(ngal);
    const double factor = std::pow((double)params._Lbox(), 1.5) / (double)params._NGRID();
    fftw_functions.set_normal_power((real_prec)std::pow(factor, -2));
    fftw_functions.shot_noise = 0;
    if (params._SN_correction())
        fftw_functions.shot_noise = (real_prec)(std::pow((double)params._Lbox(), 3) / (double)ngal);
    kvector();
    fftw_functions.power_spectrum(pk0, modes_g);
    write_power_and_modes();
}

// PowerSpectrumF.cpp:352-358: the window file carries the data shells' mode
// counts, as the reference writes it
void PowerSpectrumF::write_power_and_modes() {
    write_to_file2(file_power, kvector_data, pk0, modes_g, true);
    if (params._use_random_catalog()) write_to_file2(file_window, kvector_data, pk_w, modes_g, true);
}

void PowerSpectrumF::write_bispectrum() {
    std::ofstream out(file_bispectrum.c_str());
    if (!out) throw cmh::Error("cannot write " + file_bispectrum);
    out.precision(10);
    out.setf(std::ios::scientific);
    const int nb = (int)kvector_data_b.size();
    const size_t N = params._Nft();
    for (int a = 0; a < nb; ++a)
        for (int b = a; b < nb; ++b)
            for (int c = b; c < nb; ++c) {
                const size_t I = ((size_t)a * N + b) * N + c;
                if (modes_tri[I] == 0) continue;
                out << a << "\t" << b << "\t" << c << "\t" << kvector_data_b[a] << "\t" << kvector_data_b[b] << "\t"
                    << kvector_data_b[c] << "\t" << bispectrum[I] << "\t" << sn_bispectrum[I] << "\t" << modes_tri[I]
                    << "\n";
            }
}
