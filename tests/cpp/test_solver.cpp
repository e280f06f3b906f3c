// test_solver.cpp -- the scenarios of the reference's kernel tests
// (test/kerneltest.cpp:8-125 single component, :178-237 two components; DIM 250,
// LENGTH 100) on this project's
// host API with a HIP kernel type, plus its error paths.  Prints one JSON line
// per scenario; tests/test_cpp_api.py checks the values.
//
//   test_solver <kernel_type> [scenario]
//   test_solver io case.bin prefix      (text I/O, no device work)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <stdexcept>
#include <string>

#include "trottersuzuki.h"

#define DIM 250
#define LENGTH 100

static void report(const char *name, const std::string &kt, double e0, double e1, double n0, double n1, double xx) {
    std::printf("{\"scenario\": \"%s\", \"kernel\": \"%s\", \"E0\": %.17g, \"E\": %.17g, \"N0\": %.17g, "
                "\"N\": %.17g, \"XX\": %.17g}\n",
                name, kt.c_str(), e0, e1, n0, n1, xx);
    std::fflush(stdout);
}

static void free_particle(const std::string &kt) {
    Lattice2D grid(DIM, LENGTH, true, true);
    ExponentialState state(&grid);
    Hamiltonian ham(&grid, NULL);
    Solver solver(&grid, &state, &ham, 5.e-3, kt);
    const double e0 = solver.get_total_energy(), n0 = solver.get_squared_norm();
    solver.evolve(100);
    report("free_particle", kt, e0, solver.get_total_energy(), n0, solver.get_squared_norm(), state.get_mean_xx());
}

static void harmonic_oscillator(const std::string &kt) {
    Lattice2D grid(DIM, LENGTH);
    GaussianState state(&grid, 1.);
    HarmonicPotential pot(&grid, 1., 1.);
    Hamiltonian ham(&grid, &pot);
    Solver solver(&grid, &state, &ham, 5.e-3, kt);
    const double e0 = solver.get_total_energy(), n0 = solver.get_squared_norm();
    solver.evolve(100);
    report("harmonic_oscillator", kt, e0, solver.get_total_energy(), n0, solver.get_squared_norm(),
           state.get_mean_xx());
}

static void imaginary_harmonic_oscillator(const std::string &kt) {
    Lattice2D grid(DIM, LENGTH);
    GaussianState state(&grid, 0.5);
    HarmonicPotential pot(&grid, 1., 1.);
    Hamiltonian ham(&grid, &pot);
    Solver solver(&grid, &state, &ham, 5.e-3, kt);
    const double n0 = solver.get_squared_norm();
    solver.evolve(1000, true);
    report("imaginary_harmonic_oscillator", kt, 0.0, solver.get_total_energy(), n0, solver.get_squared_norm(),
           state.get_mean_xx());
}

static void intra_particle_interaction(const std::string &kt) {
    Lattice2D grid(DIM, LENGTH);
    GaussianState state(&grid, 1);
    HarmonicPotential pot(&grid, 1., 1.);
    Hamiltonian ham(&grid, &pot, 1., 10);
    Solver solver(&grid, &state, &ham, 1.e-3, kt);
    const double e0 = solver.get_total_energy(), n0 = solver.get_squared_norm();
    solver.evolve(1000);
    report("intra_particle_interaction", kt, e0, solver.get_total_energy(), n0, solver.get_squared_norm(),
           state.get_mean_xx());
}

static void imaginary_intra_particle_interaction(const std::string &kt) {
    Lattice2D grid(DIM, LENGTH);
    GaussianState state(&grid, 1);
    HarmonicPotential pot(&grid, 1., 1.);
    Hamiltonian ham(&grid, &pot, 1., 10);
    Solver solver(&grid, &state, &ham, 1.e-3, kt);
    const double n0 = solver.get_squared_norm();
    solver.evolve(1000, true);
    report("imaginary_intra_particle_interaction", kt, 0.0, solver.get_total_energy(), n0,
           solver.get_squared_norm(), state.get_mean_xx());
}

// The two-component scenarios (test/kerneltest.cpp:178-237) through the
// two-state Solver: state 2 starts empty, the Rabi coupling moves the norm.
static void report_pair(const char *name, const std::string &kt, double n0, double n, double n10, double n20,
                        double n1, double n2) {
    std::printf("{\"scenario\": \"%s\", \"kernel\": \"%s\", \"N0\": %.17g, \"N\": %.17g, \"N1_0\": %.17g, "
                "\"N2_0\": %.17g, \"N1\": %.17g, \"N2\": %.17g}\n",
                name, kt.c_str(), n0, n, n10, n20, n1, n2);
    std::fflush(stdout);
}

static void mixed_bec(const std::string &kt, bool imag) {
    Lattice2D grid(DIM, LENGTH);
    GaussianState state1(&grid, 1);
    State state2(&grid);
    HarmonicPotential pot(&grid, 1., 1.);
    Hamiltonian2Component ham(&grid, &pot, &pot, 1., 1., 0., 0., 0., 2. * M_PI / 10.);
    Solver solver(&grid, &state1, &state2, &ham, 1.e-3, kt);
    const double n0 = solver.get_squared_norm();
    const double n10 = state1.get_squared_norm(), n20 = state2.get_squared_norm();
    solver.evolve(imag ? 1000 : 5000, imag);
    report_pair(imag ? "imaginary_mixed_BEC" : "mixed_BEC", kt, n0, solver.get_squared_norm(), n10, n20,
                state1.get_squared_norm(), state2.get_squared_norm());
}

static void mixed_bec_real(const std::string &kt) { mixed_bec(kt, false); }
static void mixed_bec_imag(const std::string &kt) { mixed_bec(kt, true); }

// The error paths a caller can reach without a GPU fault (src/common.cpp:104-115:
// std::runtime_error with a message).
static void errors() {
    int caught = 0;
    try {
        Lattice2D grid(16, 1.);
        GaussianState st(&grid, 1.);
        Hamiltonian ham(&grid, NULL);
        Solver s(&grid, &st, &ham, 1e-3, "cpu");
    } catch (const std::runtime_error &e) {
        caught += std::strstr(e.what(), "Unknown kernel") != 0;
    }
    try {
        Lattice2D grid(16, 1., false, false, 0.5);
    } catch (const std::runtime_error &e) {
        caught += std::strstr(e.what(), "angular velocity") != 0;
    }
    try {
        Lattice2D grid(16, 1., false, false, 0., std::string("polar"));
    } catch (const std::runtime_error &e) {
        caught += std::strstr(e.what(), "not implemented") != 0;
    }
    std::printf("{\"scenario\": \"errors\", \"caught\": %d, \"expected\": 3}\n", caught);
}


// Text I/O of the state and potential (src/common.cpp:161-191, src/model.cpp:254-322,
// :342-380, :826-839): writes the state, its density and phase as text, reads the
// state back with loadtxt (halo refill on periodic axes) and a potential from a text
// file, and dumps the loaded tile and potential as doubles to <prefix>-loaded.bin.
static int run_io(const char *case_path, const std::string &prefix) {
    FILE *f = std::fopen(case_path, "rb");
    if (!f) return 2;
    int32_t hdr[10];
    double par[6];
    if (std::fread(hdr, 4, 10, f) != 10 || std::fread(par, 8, 6, f) != 6) return 2;
    Lattice2D grid(hdr[2], par[0], hdr[3], par[1], hdr[4] != 0, hdr[5] != 0);
    const size_t n = (size_t)grid.dim_x * (size_t)grid.dim_y;
    std::vector<double> re(n), im(n), v(n);
    if (std::fread(re.data(), 8, n, f) != n || std::fread(im.data(), 8, n, f) != n || std::fread(v.data(), 8, n, f) != n)
        return 2;
    std::fclose(f);
    State state(&grid, 0, re.data(), im.data());
    state.write_to_file(prefix + "-state");
    state.write_particle_density(prefix);
    state.write_phase(prefix);
    {
        FILE *pf = std::fopen((prefix + "-potential").c_str(), "w");
        for (size_t k = 0; k < n; ++k) std::fprintf(pf, "%.17g%c", v[k], (k + 1) % grid.dim_x ? ' ' : '\n');
        std::fclose(pf);
    }
    State loaded(&grid);
    std::string sp = prefix + "-state";
    loaded.loadtxt(const_cast<char *>(sp.c_str()));
    std::string pp = prefix + "-potential";
    Potential pot(&grid, const_cast<char *>(pp.c_str()));
    FILE *o = std::fopen((prefix + "-loaded.bin").c_str(), "wb");
    std::fwrite(loaded.p_real, 8, n, o);
    std::fwrite(loaded.p_imag, 8, n, o);
    for (int y = 0; y < grid.dim_y; ++y)
        for (int x = 0; x < grid.dim_x; ++x) {
            const double val = pot.get_value(x, y);
            std::fwrite(&val, 8, 1, o);
        }
    std::fclose(o);
    return 0;
}

int main(int argc, char **argv) {
    if (argc > 3 && std::strcmp(argv[1], "io") == 0) return run_io(argv[2], argv[3]);
    const std::string kt = argc > 1 ? argv[1] : "hip";
    const std::string only = argc > 2 ? argv[2] : "";
    struct { const char *name; void (*fn)(const std::string &); } all[] = {
        {"free_particle", free_particle},
        {"harmonic_oscillator", harmonic_oscillator},
        {"imaginary_harmonic_oscillator", imaginary_harmonic_oscillator},
        {"intra_particle_interaction", intra_particle_interaction},
        {"imaginary_intra_particle_interaction", imaginary_intra_particle_interaction},
        {"mixed_BEC", mixed_bec_real},
        {"imaginary_mixed_BEC", mixed_bec_imag},
    };
    try {
        if (only == "errors") {
            errors();
            return 0;
        }
        for (auto &t : all)
            if (only.empty() || only == t.name) t.fn(kt);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "test_solver: %s\n", e.what());
        return 1;
    }
    return 0;
}
