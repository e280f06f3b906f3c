// ref_driver.cpp -- drives the reference library through its public API only
// (src/trottersuzuki.h), the way the reference's examples do.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// sources (never copied into this repository) into oracle/_ref/.  Used to write
// the golden vectors under tests/golden/ and, in timing mode, as the CPU
// baseline of bench.py.
//
// Case file (little-endian, all tile arrays in the reference layout W*H incl. halo):
//   int32  magic 0x54534d49, version 1
//   int32  Nx, Ny, periodic_x, periodic_y, n_steps, imag_time, n_warmup, time_mode
//   double Lx, Ly, dt, mass, g, g_lhy
//   double psi_re[W*H], psi_im[W*H], V[W*H]
// Output file:
//   double psi_re[W*H], psi_im[W*H]
//   double energies[6]  = total, kinetic, potential, intra, LHY, squared norm (Solver getters)
//   double expect[10]   = State: norm2, X, X^2, Y, Y^2, P_x, P_x^2, P_y, P_y^2, L_z
// In time mode the driver runs n_warmup steps, then times evolve(n_steps) with
// steady_clock and prints one JSON line; no output file is written.
// `ref_driver case.bin --io prefix` exercises the text I/O instead (run_io).
//
// Version 2 case files describe a two-component system (run_pair):
//   int32  magic, version 2, Nx, Ny, periodic_x, periodic_y, n_calls, imag_time, 0, 0
//   int32  steps[n_calls]                 one Solver::evolve call each
//   double Lx, Ly, dt, mass, mass_b, g_a, g_ab, g_b, omega_r, omega_i
//   double a_re, a_im, b_re, b_im, V_a, V_b   [W*H each]
// Output: a_re, a_im, b_re, b_im [W*H each], then the Solver getters
//   total, kinetic 1, 2, potential 1, 2, intra 1, 2, inter, Rabi, squared norm 1, 2, 3
// and State::get_squared_norm of both states.
#include "trottersuzuki.h"
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

static void die(const char *m) { std::fprintf(stderr, "ref_driver: %s\n", m); std::exit(2); }

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


template <class T> static void rd(FILE *f, T *p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) die("short read");
}

// Two-component system through Hamiltonian2Component and the two-state Solver
// (src/trottersuzuki.h:459-497, src/solver.cpp:43-60,150-242).
static int run_pair(FILE *f, const int32_t *hdr, const char *out_path) {
    const int Nx = hdr[2], Ny = hdr[3], px = hdr[4], py = hdr[5], n_calls = hdr[6], imag = hdr[7];
    std::vector<int32_t> calls(n_calls > 0 ? n_calls : 0);
    rd(f, calls.data(), calls.size());
    double par[10];
    rd(f, par, 10);
    Lattice2D grid(Nx, par[0], Ny, par[1], px != 0, py != 0);
    const size_t n = (size_t)grid.dim_x * (size_t)grid.dim_y;
    std::vector<double> ar(n), ai(n), br(n), bi(n), va(n), vb(n);
    rd(f, ar.data(), n);
    rd(f, ai.data(), n);
    rd(f, br.data(), n);
    rd(f, bi.data(), n);
    rd(f, va.data(), n);
    rd(f, vb.data(), n);
    std::fclose(f);
    State sa(&grid, 0, ar.data(), ai.data());
    State sb(&grid, 0, br.data(), bi.data());
    Potential pa(&grid, va.data()), pb(&grid, vb.data());
    Hamiltonian2Component ham(&grid, &pa, &pb, par[3], par[4], par[5], par[6], par[7], par[8], par[9]);
    Solver solver(&grid, &sa, &sb, &ham, par[2], "cpu");
    for (int32_t c : calls) solver.evolve(c, imag != 0);
    double e[14];
    e[0] = solver.get_total_energy();
    e[1] = solver.get_kinetic_energy(1);
    e[2] = solver.get_kinetic_energy(2);
    e[3] = solver.get_potential_energy(1);
    e[4] = solver.get_potential_energy(2);
    e[5] = solver.get_intra_species_energy(1);
    e[6] = solver.get_intra_species_energy(2);
    e[7] = solver.get_inter_species_energy();
    e[8] = solver.get_rabi_energy();
    e[9] = solver.get_squared_norm(1);
    e[10] = solver.get_squared_norm(2);
    e[11] = solver.get_squared_norm(3);
    e[12] = sa.get_squared_norm();
    e[13] = sb.get_squared_norm();
    FILE *o = std::fopen(out_path, "wb");
    if (!o) die("cannot open output");
    std::fwrite(ar.data(), 8, n, o);
    std::fwrite(ai.data(), 8, n, o);
    std::fwrite(br.data(), 8, n, o);
    std::fwrite(bi.data(), 8, n, o);
    std::fwrite(e, 8, 14, o);
    std::fclose(o);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) die("usage: ref_driver case.bin [out.bin | --io prefix]");
    if (argc > 3 && std::strcmp(argv[2], "--io") == 0) return run_io(argv[1], argv[3]);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) die("cannot open case file");
    int32_t hdr[10];
    rd(f, hdr, 10);
    if (hdr[0] == 0x54534d49 && hdr[1] == 2) {
        if (argc < 3) die("missing output path");
        return run_pair(f, hdr, argv[2]);
    }
    if (hdr[0] != 0x54534d49 || hdr[1] != 1) die("bad case header");
    const int Nx = hdr[2], Ny = hdr[3], px = hdr[4], py = hdr[5];
    const int n_steps = hdr[6], imag = hdr[7], n_warm = hdr[8], time_mode = hdr[9];
    double par[6];
    rd(f, par, 6);
    const double Lx = par[0], Ly = par[1], dt = par[2], mass = par[3], g = par[4], g_lhy = par[5];

    Lattice2D grid(Nx, Lx, Ny, Ly, px != 0, py != 0);
    const size_t n = (size_t)grid.dim_x * (size_t)grid.dim_y;
    std::vector<double> re(n), im(n), v(n);
    rd(f, re.data(), n);
    rd(f, im.data(), n);
    rd(f, v.data(), n);
    std::fclose(f);

    State state(&grid, 0, re.data(), im.data());
    Potential pot(&grid, v.data());
    Hamiltonian ham(&grid, &pot, mass, g, g_lhy);
    Solver solver(&grid, &state, &ham, dt, "cpu");

    if (time_mode) {
        if (n_warm > 0) solver.evolve(n_warm, imag != 0);
        auto t0 = std::chrono::steady_clock::now();
        solver.evolve(n_steps, imag != 0);
        auto t1 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t1 - t0).count();
        int threads = 1;
#ifdef _OPENMP
        threads = omp_get_max_threads();
#endif
        std::printf("{\"seconds\": %.9g, \"steps\": %d, \"dim_x\": %d, \"dim_y\": %d, \"threads\": %d}\n",
                    s, n_steps, grid.dim_x, grid.dim_y, threads);
        return 0;
    }

    solver.evolve(n_steps, imag != 0);
    double e[6];
    e[0] = solver.get_total_energy();
    e[1] = solver.get_kinetic_energy();
    e[2] = solver.get_potential_energy();
    e[3] = solver.get_intra_species_energy();
    e[4] = solver.get_LeeHuangYang_energy();
    e[5] = solver.get_squared_norm();
    double x[10];
    x[0] = state.get_squared_norm();
    x[1] = state.get_mean_x();
    x[2] = state.get_mean_xx();
    x[3] = state.get_mean_y();
    x[4] = state.get_mean_yy();
    x[5] = state.get_mean_px();
    x[6] = state.get_mean_pxpx();
    x[7] = state.get_mean_py();
    x[8] = state.get_mean_pypy();
    x[9] = state.get_mean_angular_momentum();
    if (argc < 3) die("missing output path");
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) die("cannot open output");
    std::fwrite(re.data(), sizeof(double), n, o);
    std::fwrite(im.data(), sizeof(double), n, o);
    std::fwrite(e, sizeof(double), 6, o);
    std::fwrite(x, sizeof(double), 10, o);
    std::fclose(o);
    return 0;
}
