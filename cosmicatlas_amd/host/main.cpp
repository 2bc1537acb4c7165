// cosmicatlas CLI for the mesh-measurement path (Main/cosmicatlass.cpp:38-165):
//   -p <ini>  measure P(k) (Input_type = catalog | density_grid | delta_grid),
//             plus the T-web when compute_tweb = true
//   -i <ini>  show the parameters and derived constants
//   -h        usage
// Errors print a message and exit(1), like the reference.
#include <cstdio>
#include <cstring>
#include <iostream>

#include "cosmicatlas_host.h"

static void usage(const char *prog) {
    std::printf("Usage: %s -p parameter_file.ini   measure the power spectrum (and T-web)\n"
                "       %s -i parameter_file.ini   show the parameters\n"
                "       %s -h                      this message\n",
                prog, prog, prog);
}

int main(int argc, char **argv) {
    if (argc == 1) {
        usage(argv[0]);
        return 1;
    }
    try {
        if (!std::strcmp(argv[1], "-h")) {
            usage(argv[0]);
            return 0;
        }
        if (argc < 3) {
            usage(argv[0]);
            return 1;
        }
        const std::string opt = argv[1], ini = argv[2];
        if (opt == "-i") {
            Params params(ini);
            params.show_params();
            return 0;
        }
        if (opt == "-p") {
            Params params(ini);
            if (params._measure_cross())
                throw cmh::Error("measure_cross: cross power spectra are not part of this build");
            PowerSpectrumF cPSF(params);
            if (params._input_type() == "density_grid" || params._input_type() == "delta_grid")
                cPSF.compute_power_spectrum_grid();
            else if (params._input_type() == "catalog")
                cPSF.compute_power_spectrum(true, false);
            else
                throw cmh::Error("Unknown Input_type " + params._input_type());
            std::printf("Power spectrum written to %s\n", cPSF.file_power.c_str());
            return 0;
        }
        usage(argv[0]);
        return 1;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl << "Cosmicatlass stops here." << std::endl;
        return 1;
    }
}
