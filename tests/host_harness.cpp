// Test driver of the C++ host layer's free functions (reference names and
// argument meanings): reads float32 inputs from files, calls the function,
// writes float32 outputs.  Used by tests/test_host_functions.py against the
// oracle.
//   getdensity SCHEME N L weightmass pos.bin(x[n] y[n] z[n] m[n]) out.bin [init.bin]
//                               (init.bin: the N^3 mesh delta holds on entry)
//   r2c N in.bin out.bin        do_fftw_r2c: N^3 reals -> N*N*(N/2+1) complex
//   c2r N in.bin out.bin        do_fftw_c2r: half spectrum -> N^3 reals
//   poisson N L in.bin out.bin  PoissonSolver
//   eigen N L phi.bin out.bin   EigenValuesTweb (FD branch): l1|l2|l3
//   nobjects in.bin out.bin     get_nobjects
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../cosmicatlas_amd/host/cosmicatlas_host.h"

static std::vector<real_prec> load(const char *f) {
    std::ifstream in(f, std::ios::binary | std::ios::ate);
    if (!in) throw cmh::Error(std::string("cannot read ") + f);
    const size_t bytes = (size_t)in.tellg();
    std::vector<real_prec> v(bytes / sizeof(real_prec));
    in.seekg(0);
    in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)bytes);
    return v;
}

static void save(const char *f, const void *p, size_t bytes) {
    std::ofstream out(f, std::ios::binary);
    if (!out) throw cmh::Error(std::string("cannot write ") + f);
    out.write(reinterpret_cast<const char *>(p), (std::streamsize)bytes);
}

int main(int argc, char **argv) {
    try {
        if (argc < 2) throw cmh::Error("usage: host_harness COMMAND ...");
        const std::string cmd = argv[1];
        if (cmd == "getdensity" && (argc == 8 || argc == 9)) {
            const std::string sch = argv[2];
            const ULONG N = std::strtoul(argv[3], nullptr, 10);
            const real_prec L = (real_prec)std::atof(argv[4]);
            const bool wm = std::atoi(argv[5]) != 0;
            std::vector<real_prec> all = load(argv[6]);
            const size_t n = all.size() / 4;
            std::vector<real_prec> x(all.begin(), all.begin() + n), y(all.begin() + n, all.begin() + 2 * n),
                z(all.begin() + 2 * n, all.begin() + 3 * n), m(all.begin() + 3 * n, all.begin() + 4 * n), delta;
            if (argc == 9) delta = load(argv[8]);
            const real_prec d = L / (real_prec)N;
            auto f = sch == "NGP" ? getDensity_NGP : sch == "CIC" ? getDensity_CIC : getDensity_TSC;
            f(N, N, N, L, L, L, d, d, d, 0, 0, 0, x, y, z, m, delta, wm);
            save(argv[7], delta.data(), sizeof(real_prec) * delta.size());
        } else if (cmd == "r2c" && argc == 5) {
            const ULONG N = std::strtoul(argv[2], nullptr, 10);
            std::vector<real_prec> in = load(argv[3]);
            std::vector<real_prec> out(2 * (size_t)N * N * (N / 2 + 1));
            do_fftw_r2c(N, in, reinterpret_cast<complex_prec *>(out.data()));
            save(argv[4], out.data(), sizeof(real_prec) * out.size());
        } else if (cmd == "c2r" && argc == 5) {
            const int N = std::atoi(argv[2]);
            std::vector<real_prec> in = load(argv[3]), out;
            do_fftw_c2r(N, reinterpret_cast<complex_prec *>(in.data()), out);
            save(argv[4], out.data(), sizeof(real_prec) * out.size());
        } else if (cmd == "poisson" && argc == 6) {
            const ULONG N = std::strtoul(argv[2], nullptr, 10);
            const real_prec L = (real_prec)std::atof(argv[3]);
            std::vector<real_prec> in = load(argv[4]), out;
            PoissonSolver(L, N, in, out);
            save(argv[5], out.data(), sizeof(real_prec) * out.size());
        } else if (cmd == "eigen" && argc == 6) {
            const ULONG N = std::strtoul(argv[2], nullptr, 10);
            const real_prec L = (real_prec)std::atof(argv[3]);
            std::vector<real_prec> phi = load(argv[4]), delta, l1, l2, l3;
            EigenValuesTweb(N, L, delta, phi, l1, l2, l3);
            std::vector<real_prec> all;
            all.insert(all.end(), l1.begin(), l1.end());
            all.insert(all.end(), l2.begin(), l2.end());
            all.insert(all.end(), l3.begin(), l3.end());
            save(argv[5], all.data(), sizeof(real_prec) * all.size());
        } else if (cmd == "nobjects" && argc == 4) {
            std::vector<real_prec> in = load(argv[2]);
            const real_prec s = get_nobjects(in);
            save(argv[3], &s, sizeof s);
        } else {
            throw cmh::Error("unknown command or wrong argument count: " + cmd);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
