// Host-side C++ interface of the mesh-measurement path, keeping the names,
// argument meaning and data layouts of the COSMICATLAS classes and free
// functions it replaces.  Everything computes on the GPU through the C ABI in
// include/cosmo_mesh.h; this layer only holds host vectors, parameters and
// files.  Errors throw cmh::Error (the CLI turns them into the reference's
// message-and-exit(1)).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cosmo_mesh.h"

typedef float real_prec;              // def.h:491,580 (float build)
typedef real_prec complex_prec[2];    // fftwf_complex (def.h:587)
typedef unsigned long ULONG;
#ifndef REAL
#define REAL 0
#define IMAG 1
#endif
// Cosmic-web labels (def.h:644-647)
#define I_KNOT 1
#define I_FILAMENT 2
#define I_SHEET 3
#define I_VOID 4

namespace cmh {
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// Process-wide library context on device 0 (created on first use).
cm_ctx *ctx();
void check(cm_status s, const char *what);
}  // namespace cmh

// --------------------------------------------------------------------- Params
// .ini reader (Source/Params.cpp:452-481: "name = value", first token after '='
// is the value, '#' lines skipped, empty value keeps the default) and the
// derived constants of derived_pars (Source/Params.cpp:4079-4115).  Keys the
// reference fixes at compile time (def.h) are run-time keys here:
//   shell_rule = exact | float   (float = FftwFunctions.cpp:1097 float32 index)
//   compute_tweb = true | false  (run Cwclass on the measured mesh, -p only)
class Params {
   public:
    Params();
    explicit Params(const std::string &par_file);
    void init_pars();
    void read_pars(const std::string &file);
    void derived_pars();
    void show_params() const;
    std::string get(const std::string &key, const std::string &def = "") const;

    ULONG _Nft() const { return Nft; }
    real_prec _Lbox() const { return Lbox; }
    ULONG _NGRID() const { return NGRID; }
    ULONG _NGRID_h() const { return NGRID_h; }
    std::string _mass_assignment_scheme() const { return mass_assignment_scheme; }
    int _mass_assignment() const { return mass_assignment; }
    bool _MAS_correction() const { return MAS_correction; }
    std::string _type_of_binning() const { return type_of_binning; }
    real_prec _DeltaKmin() const { return DeltaKmin; }
    int _ndel_data() const { return ndel_data; }
    int _N_log_bins() const { return N_log_bins; }
    bool _SN_correction() const { return SN_correction; }
    real_prec _lambdath() const { return lambdath; }
    real_prec _smscale() const { return smscale; }
    std::string _statistics() const { return statistics; }
    std::string _input_type() const { return input_type; }
    std::string _Output_directory() const { return Output_directory; }
    std::string _file_power() const { return file_power; }
    std::string _file_bispectrum() const { return file_bispectrum; }
    double _kmax_bk() const { return kmax_bk; }
    double _kmin_bk() const { return kmin_bk; }
    std::string _Name_survey() const { return Name_survey; }
    std::string _delta_grid_file() const { return delta_grid_file; }
    std::string _file_catalogue() const { return Input_dir_cat + file_catalogue; }
    bool _use_random_catalog() const { return use_random_catalog; }
    std::string _file_random() const { return Input_dir_cat + file_random; }
    std::string _file_window() const { return file_window; }
    // random-catalogue columns; an unset coordinate column (-1) takes the
    // tracer catalogue's
    int _i_coord1_r() const { return i_coord1_r >= 0 ? i_coord1_r : i_coord1_g; }
    int _i_coord2_r() const { return i_coord2_r >= 0 ? i_coord2_r : i_coord2_g; }
    int _i_coord3_r() const { return i_coord3_r >= 0 ? i_coord3_r : i_coord3_g; }
    int _i_weight1_r() const { return i_weight1_r; }
    int _i_mean_density_g() const { return i_mean_density_g; }
    int _i_mean_density_r() const { return i_mean_density_r; }
    bool _FKP_weight() const { return FKP_weight; }
    real_prec _Pest() const { return Pest; }
    bool _measure_cross() const { return measure_cross; }
    int _i_coord1_g() const { return i_coord1_g; }
    int _i_coord2_g() const { return i_coord2_g; }
    int _i_coord3_g() const { return i_coord3_g; }
    int _i_weight1_g() const { return i_weight1_g; }
    int _i_v1_g() const { return i_v1_g; }
    int _i_v2_g() const { return i_v2_g; }
    int _i_v3_g() const { return i_v3_g; }
    std::string _vel_units_g() const { return vel_units_g; }
    bool _redshift_space_coords_g() const { return redshift_space_coords_g; }
    double _redshift() const { return redshift; }
    real_prec _Xoffset() const { return Xoffset; }
    real_prec _Yoffset() const { return Yoffset; }
    real_prec _Zoffset() const { return Zoffset; }
    double _d_delta_x() const { return delta_x; }
    double _d_deltak_x() const { return deltak_x; }
    double _d_deltak_0() const { return deltak_0; }
    double _d_kmin() const { return kmin; }
    double _d_kmax() const { return kmax; }
    double _d_DeltaK_data() const { return DeltaK_data; }
    int _d_Nnp_data() const { return Nnp_data; }
    double _d_Deltal() const { return Deltal; }
    bool _float_shell_rule() const { return shell_rule == "float"; }
    bool _compute_tweb() const { return compute_tweb; }
    std::string _file_bin_x_coord() const { return file_bin_x_coord; }
    std::string _file_bin_y_coord() const { return file_bin_y_coord; }
    std::string _file_bin_z_coord() const { return file_bin_z_coord; }
    ULONG _N_lines_binary() const { return N_lines_binary; }
    int _sys_of_coord_g() const { return sys_of_coord_g; }
    std::string _angles_units_g() const { return angles_units_g; }
    double _om_matter() const { return om_matter; }
    double _om_radiation() const { return om_radiation; }
    double _om_vac() const { return om_vac; }
    double _om_k() const { return om_k; }
    double _wde_eos() const { return wde_eos; }
    double _Hubble() const { return Hubble; }
    int _Nbins_redshift() const { return Nbins_redshift; }
    double _redshift_min_sample() const { return redshift_min_sample; }
    double _redshift_max_sample() const { return redshift_max_sample; }

   private:
    std::map<std::string, std::string> raw;
    ULONG Nft = 1, NGRID = 1, NGRID_h = 1;
    real_prec Lbox = 100;
    std::string mass_assignment_scheme = "NGP";
    int mass_assignment = 0;
    bool MAS_correction = false;
    std::string type_of_binning = "linear";
    real_prec DeltaKmin = 0;
    int ndel_data = 1, N_log_bins = 10;
    bool SN_correction = false;
    real_prec lambdath = 0, smscale = 0;
    std::string statistics = "Pk_fkp", input_type = "catalog", Output_directory = "../output";
    std::string file_power = "power", Name_survey = "survey", delta_grid_file = "delta_file";
    // bispectrum output stem and shell range (Params.cpp:361-366)
    std::string file_bispectrum = "bispectrum";
    double kmax_bk = 0.2, kmin_bk = 0.1;
    std::string file_catalogue = "cat.dat", Input_dir_cat = "";
    bool use_random_catalog = false, measure_cross = false;
    // random catalogue and FKP weights (Params.cpp:228,262-266,285,341-342,357)
    std::string file_random = "random", file_window = "window";
    int i_coord1_r = -1, i_coord2_r = -1, i_coord3_r = -1, i_weight1_r = -1;
    int i_mean_density_g = -1, i_mean_density_r = -1;
    bool FKP_weight = false;
    real_prec Pest = 1000;
    int i_coord1_g = 0, i_coord2_g = 1, i_coord3_g = 2, i_weight1_g = -1;
    // velocity columns (-1: absent), their units and whether the coordinates
    // already hold the redshift-space displacement (Params.cpp:170,215-227)
    int i_v1_g = -1, i_v2_g = -1, i_v3_g = -1;
    std::string vel_units_g = "kmps";
    bool redshift_space_coords_g = false;
    double redshift = 0.0;  // the snapshot's redshift (Params.cpp:127,567-570)
    // binary SoA catalogues (Catalog::read_catalog_bin, Source/Catalog.cpp:560):
    // one headerless real_prec file per coordinate
    std::string file_bin_x_coord, file_bin_y_coord, file_bin_z_coord;
    ULONG N_lines_binary = 0;
    // coordinate system of the catalogue: 0 Cartesian, 2 equatorial (RA, Dec,
    // comoving distance) converted by equatorial_to_cartesian
    // (Source/CoordinateSystem.cpp:12-18); angles in degrees ("D") or radians ("R")
    int sys_of_coord_g = 0;
    std::string angles_units_g = "D";
    // background cosmology for the distance-redshift relation: the .ini keys of
    // Source/Params.cpp:3168-3219, defaults of Cosmo_parameters_PLANCK
    // (Headers/cosmo_parameters.h:37-50); H in km/s per Mpc/h
    double om_matter = 0.3089;
    double om_radiation = 1.3157e-5 * (1.0 + 0.2271 * 3.046);
    double om_k = 0.0;
    double om_vac = 1.0 - 0.3089 - 0.0 - 1.3157e-5 * (1.0 + 0.2271 * 3.046);
    double wde_eos = -1.0;
    double Hubble = 100.0;
    // r(z) table nodes (Params.cpp:348-350)
    int Nbins_redshift = 10;
    double redshift_min_sample = 0.0, redshift_max_sample = 1.0;
    real_prec Xoffset = 0, Yoffset = 0, Zoffset = 0;
    double delta_x = 0, deltak_x = 0, deltak_0 = 0, kmin = 0, kmax = 0, DeltaK_data = 0, Deltal = 0;
    int Nnp_data = 0;
    std::string shell_rule = "exact";
    bool compute_tweb = false;
};

// ------------------------------------------------------------ free functions
// Source/Miscelanious.cpp:65-147,454-557,693-815 (array form, SoA positions)
void getDensity_NGP(ULONG N1, ULONG N2, ULONG N3, real_prec L1, real_prec L2, real_prec L3, real_prec d1,
                    real_prec d2, real_prec d3, real_prec min1, real_prec min2, real_prec min3,
                    const std::vector<real_prec> &xp, const std::vector<real_prec> &yp,
                    const std::vector<real_prec> &zp, const std::vector<real_prec> &Par_mass,
                    std::vector<real_prec> &delta, bool weightmass);
void getDensity_CIC(ULONG N1, ULONG N2, ULONG N3, real_prec L1, real_prec L2, real_prec L3, real_prec d1,
                    real_prec d2, real_prec d3, real_prec min1, real_prec min2, real_prec min3,
                    const std::vector<real_prec> &xp, const std::vector<real_prec> &yp,
                    const std::vector<real_prec> &zp, const std::vector<real_prec> &Par_mass,
                    std::vector<real_prec> &delta, bool weightmass);
void getDensity_TSC(ULONG N1, ULONG N2, ULONG N3, real_prec L1, real_prec L2, real_prec L3, real_prec d1,
                    real_prec d2, real_prec d3, real_prec min1, real_prec min2, real_prec min3,
                    const std::vector<real_prec> &xp, const std::vector<real_prec> &yp,
                    const std::vector<real_prec> &zp, const std::vector<real_prec> &Par_mass,
                    std::vector<real_prec> &delta, bool weightmass);
// Catalogue-to-mesh for all tracers (grid_assignment_* per tracer,
// Source/Miscelanious.cpp:2597-2868 as called from Catalog.cpp:14186-14211).
void grid_assignment(const Params &params, const std::vector<real_prec> &xyz /* AoS */,
                     const std::vector<real_prec> &w /* empty = 1 */, std::vector<real_prec> &field);
// Source/NumericalMethods.cpp:1718-1753 / :1818-1856
void do_fftw_r2c(ULONG Nft, std::vector<real_prec> &in, complex_prec *out);
void do_fftw_c2r(int Nft, complex_prec *in, std::vector<real_prec> &out);
// Source/Miscelanious.cpp:1189-1274
void PoissonSolver(real_prec Lbox, ULONG Nft, std::vector<real_prec> &in, std::vector<real_prec> &out);
// Source/Miscelanious.cpp:1668-1769 (FD branch)
void EigenValuesTweb(ULONG Nft, real_prec L1, std::vector<real_prec> &delta, std::vector<real_prec> &phi,
                     std::vector<real_prec> &out1, std::vector<real_prec> &out2, std::vector<real_prec> &out3);
// Source/Miscelanious.cpp:1421-1426 (summed in FP64 here)
real_prec get_nobjects(const std::vector<real_prec> &density_field);

// --------------------------------------------------------------- FftwFunctions
// Source/FftwFunctions.cpp:28-406,958-1875 (box estimator).
class FftwFunctions {
   public:
    FftwFunctions() = default;
    explicit FftwFunctions(const Params &p) : params(p) {}
    ~FftwFunctions();
    void set_params(const Params &p) { params = p; }
    void resize_fftw_vectors();
    void free_fftw_vectors();
    void set_n_gal(real_prec n) { n_gal = n; }
    void set_w_g(real_prec w) { w_g = w; }
    void set_s_g(real_prec s) { s_g = s; }
    void set_normal_power(real_prec v) { normal_power = v; }
    // random catalogue sums (Catalog::get_interpolated_density_field,
    // Source/Catalog.cpp:14153-14171): W_r, S_r = sum w^2, normal_p = sum nbar w^2
    void set_n_ran(real_prec n) { n_ran = n; }
    void set_w_r(real_prec w) { w_r = w; }
    void set_s_r(real_prec s) { s_r = s; }
    void set_normal_p(real_prec v) { normal_p = v; }
    void raw_sampling(real_prec vol);
    void get_parameters_estimator(bool verbose);
    void get_fluctuation();
    // monopole/quadrupole/hexadecapole of the box FKP estimator (r2c inside)
    void get_power_spectrum_fkp(std::vector<real_prec> &power_g0, std::vector<real_prec> &power_g2,
                                std::vector<real_prec> &power_g4, std::vector<int> &mod_g);
    // the same with the window W(k) of the random catalogue (power_r, its modes
    // mod_r) when use_random_catalog (FftwFunctions.cpp:1248-1262, 1635-1638)
    void get_power_spectrum_fkp(std::vector<real_prec> &power_g0, std::vector<real_prec> &power_g2,
                                std::vector<real_prec> &power_g4, std::vector<real_prec> &power_r,
                                std::vector<int> &mod_g, std::vector<int> &mod_r);
    // grid-input monopole (r2c inside); the reference's D2 defect at
    // FftwFunctions.cpp:1853 is not reproduced
    void power_spectrum(std::vector<real_prec> &powerk_g0, std::vector<int> &mod_g);
    // direct bispectrum (FftwFunctions.cpp:3105-3350; r2c inside): Nft^3 arrays at
    // index_3d(a, b, c) = (a Nft + b) Nft + c, a <= b <= c over the Nnp_data bins:
    // bispect = B minus the Poisson shot noise (SN_correction), sn_bispect = the raw
    // sum of |d1|^2+|d2|^2+|d3|^2, mod = ordered pairs (k1, k2)
    void get_bispectrum_fkp(char d_w, std::vector<real_prec> &bispect, std::vector<real_prec> &sn_bispect,
                            std::vector<int> &mod);
    // shell-transform bispectrum (FftwFunctions.cpp:3358-3718, 4878): shells of
    // width DeltaK_data from kmin up to kmax_bk, triangle loop from kmin_bk; writes
    // "i j k B SN ntri" rows to file; pk receives the shell power of the first shells
    void get_bispectrum_fkp_fast(std::vector<real_prec> &pk, std::vector<real_prec> &bispect, std::vector<int> &mod,
                                 const std::string &file);
    real_prec _normal_power() const { return normal_power; }
    real_prec _shot_noise() const { return shot_noise; }
    real_prec _alpha() const { return alpha; }
    real_prec _normal_window() const { return normal_window; }

    std::vector<real_prec> data_g;    // FftwFunctions.h:1179
    std::vector<real_prec> data_r;    // the random catalogue's mesh
    complex_prec *data_out_g = nullptr;
    real_prec shot_noise = 0;

   private:
    void shells(const std::vector<real_prec> &mesh, bool window, std::vector<real_prec> &p0,
                std::vector<real_prec> *p2, std::vector<real_prec> *p4, std::vector<int> &mod);
    Params params;
    real_prec n_gal = 0, w_g = 0, s_g = 0, normal_power = 1, alpha = 1, normal_p = 1, s_r = 0;
    real_prec n_ran = 0, w_r = 0, normal_window = 1;
};

// --------------------------------------------------------------------- Cwclass
// Source/Cwclass.cpp:235-894 (T-web), Cwclass.h:138,171-183.
class Cwclass {
   public:
    Cwclass() = default;
    explicit Cwclass(const Params &p) : params(p) {}
    void set_params(const Params &p) { params = p; }
    // delta = overdensity; smoothing R = smscale (kernelcomp, applied here)
    void get_CWC(std::vector<real_prec> &delta);
    // combos of classes requested by the parameter file (t_CWT_used): codes
    // 1..4, 12, 23, 34, 123, 234 (Source/Cwclass.cpp:1925-2030)
    std::vector<int> cwt_used;
    bool get_cell_classified(int sua, ULONG ig) const;
    int get_Tclassification(ULONG ig) const;
    std::vector<int> CWClass;
    std::vector<real_prec> lambda1, lambda2, lambda3;
    real_prec volume_knots = 0, volume_filaments = 0, volume_sheets = 0, volume_voids = 0;
    int64_t n_marginal = 0;   // cells with an eigenvalue within margin_tol of lambdath
    double margin_tol = 0.0;  // the tolerance applied (float32 rounding of phi)
    // tidal tensor of get_CWC: the FD Hessian of phi (the reference's default
    // branch) or T_ab = k_a k_b / k^2 delta_k (its _USE_GFFT_EIGENV_ branch,
    // a compile-time choice there, a run-time one here)
    bool spectral_hessian = false;

   private:
    Params params;
};

// --------------------------------------------------------------- PowerSpectrumF
// Source/PowerSpectrumF.cpp:15-46,350-355,741-1363,4188-4290.
class PowerSpectrumF {
   public:
    explicit PowerSpectrumF(const Params &p);
    void set_output_filenames();
    void compute_power_spectrum(bool verbose, bool mcut);
    // use_random_catalog: tracer and random catalogues onto two meshes with
    // their FKP weights and sums, the random branch of the estimator, P(k)
    // and the window W(k) (PowerSpectrumF.cpp:741-1363)
    void compute_power_spectrum_randoms(bool verbose);
    void compute_power_spectrum_grid();
    void write_power_and_modes();
    void write_bispectrum();
    std::vector<real_prec> kvector_data, pk0, pk2, pk4, pk_w;
    std::vector<int> modes_g, modes_r;
    std::string file_power, file_cwc, file_bispectrum, file_window;
    std::vector<real_prec> kvector_data_b, bispectrum, sn_bispectrum;
    std::vector<int> modes_tri;
    FftwFunctions fftw_functions;
    Cwclass cwclass;

   private:
    void kvector();
    Params params;
};

// Raw float32 meshes (FileOutput.cpp:727-753,858-912) and the k P N text file
// (FileOutput.cpp:490-511: row 0 dropped when P(k_0) < 1).
void read_array(const std::string &fname, std::vector<real_prec> &out);
void write_array(const std::string &fname, const std::vector<real_prec> &in);
void write_to_file2(const std::string &fname, const std::vector<real_prec> &kve, const std::vector<real_prec> &pk,
                    const std::vector<int> &si, bool cero);
// Comoving distance in Mpc/h, c * int_0^z dz'/H(z') (Cosmology::comoving_distance,
// Source/CosmologicalFunctions.cpp:51-53), and the r(z) table on Nbins_redshift
// equally spaced redshifts from 0 to redshift_max_sample (Catalog::get_cosmo,
// Source/Catalog.cpp:16-37).
double comoving_distance(const Params &p, double z);
void redshift_table(const Params &p, std::vector<double> &zc, std::vector<double> &rc);
// Line-of-sight velocity -> displacement in Mpc/h for redshift-space
// coordinates (Source/Catalog.cpp:14088-14099): (1+z)/H(z) for km/s.
double rsd_conversion_factor(const Params &p);
// ASCII catalogue: whitespace-separated columns (Catalog::read_catalog); a
// velocity column i_v3_g moves objects along the third axis into redshift space.
void read_catalog(const Params &p, std::vector<real_prec> &xyz, std::vector<real_prec> &w);
// The same reader for the random catalogue (Random_file, i_coord*_r,
// i_weight1_r) and either catalogue's mean-density column (i_mean_density_g or
// _r; nb stays empty without one).  Coordinates go through the tracer's
// coordinate system (sys_of_coord_g).
void read_catalog_columns(const Params &p, const std::string &file, const int c[3], int cw, int cnb,
                          std::vector<real_prec> &xyz, std::vector<real_prec> &w, std::vector<real_prec> &nb);
