// test_solver.cpp -- the single-component scenarios of the reference's kernel
// tests (test/kerneltest.cpp:8-125: DIM 250, LENGTH 100) on this project's
// host API with a HIP kernel type, plus its error paths.  Prints one JSON line
// per scenario; tests/test_cpp_api.py checks the values.
//
//   test_solver <kernel_type> [scenario]
#include <cstdio>
#include <cstring>
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

int main(int argc, char **argv) {
    const std::string kt = argc > 1 ? argv[1] : "hip";
    const std::string only = argc > 2 ? argv[2] : "";
    struct { const char *name; void (*fn)(const std::string &); } all[] = {
        {"free_particle", free_particle},
        {"harmonic_oscillator", harmonic_oscillator},
        {"imaginary_harmonic_oscillator", imaginary_harmonic_oscillator},
        {"intra_particle_interaction", intra_particle_interaction},
        {"imaginary_intra_particle_interaction", imaginary_intra_particle_interaction},
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
