// dropin_check.cpp -- HipKernel inside the reference's own object model.
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref/ from the reference sources,
// run by tests/test_cpp_api.py on the GPU host).  Compiled against the
// reference's trottersuzuki.h, it builds the reference's Lattice2D, State,
// Potential and Hamiltonian from a case file (the format of ref_driver.cpp),
// constructs HipKernel with the arguments the reference's kernel factory passes
// to CPUBlock (src/solver.cpp:124), drives it with the reference's per-step
// protocol (src/solver.cpp:186-239) and compares the result with the reference
// Solver's own "cpu" run of the same case.
//
// A version-2 case file (two components, the format of ref_driver.cpp) builds
// Hamiltonian2Component and two states instead and drives HipPairKernel with the
// two-component protocol of Solver::evolve (src/solver.cpp:172-230), one
// protocol run per entry of the case's call list, against the reference
// two-state Solver.
//
//   dropin_check case.bin [precision [calls [update_at]]]   ->  one JSON line
//
// calls ("7,5"): single component, one Solver::evolve call per entry instead of
// the case's step count, on both sides.  update_at: before that step (counted
// over all calls) the kernel gets ITrotterKernel::update_potential with the same
// table, as Solver::evolve does when Potential::update reports a change
// (src/solver.cpp:187-192).  The line reports the number of two-step passes the
// kernel ran (fused_launches).
#include <trottersuzuki.h>
#include <hip_kernel.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static void die(const char *m) { std::fprintf(stderr, "dropin_check: %s\n", m); std::exit(2); }

static void exp_table(Potential &pot, Lattice2D &grid, double dt, bool imag, std::vector<double> &re,
                      std::vector<double> &im) {
    for (int y = 0; y < grid.dim_y; ++y)
        for (int x = 0; x < grid.dim_x; ++x) {
            const double p = pot.get_value(x, y);
            const std::complex<double> e = imag ? std::exp(std::complex<double>(-dt * p, 0.))
                                                : std::exp(std::complex<double>(0., -dt * p));
            re[(size_t)y * grid.dim_x + x] = e.real();
            im[(size_t)y * grid.dim_x + x] = e.imag();
        }
}

static double rel_l2(const std::vector<double> &ar, const std::vector<double> &ai, const std::vector<double> &br,
                     const std::vector<double> &bi) {
    double num = 0, den = 0;
    for (size_t i = 0; i < ar.size(); ++i) {
        const double dr = ar[i] - br[i], di = ai[i] - bi[i];
        num += dr * dr + di * di;
        den += br[i] * br[i] + bi[i] * bi[i];
    }
    return std::sqrt(num / den);
}

static int run_pair(FILE *f, const int32_t *h, int precision) {
    std::vector<int32_t> calls(h[6] > 0 ? h[6] : 0);
    double par[10];
    if (std::fread(calls.data(), 4, calls.size(), f) != calls.size() || std::fread(par, 8, 10, f) != 10)
        die("short header");
    Lattice2D grid(h[2], par[0], h[3], par[1], h[4] != 0, h[5] != 0);
    const size_t n = (size_t)grid.dim_x * grid.dim_y;
    std::vector<double> ar(n), ai(n), br(n), bi(n), va(n), vb(n);
    for (std::vector<double> *a : {&ar, &ai, &br, &bi, &va, &vb})
        if (std::fread(a->data(), 8, n, f) != n) die("short data");
    std::fclose(f);
    const bool imag = h[7] != 0;
    const double dt = par[2];
    Potential pa(&grid, va.data()), pb(&grid, vb.data());
    Hamiltonian2Component ham(&grid, &pa, &pb, par[3], par[4], par[5], par[6], par[7], par[8], par[9]);

    // reference two-state Solver on copies
    std::vector<double> car(ar), cai(ai), cbr(br), cbi(bi);
    State ca(&grid, 0, car.data(), cai.data()), cb(&grid, 0, cbr.data(), cbi.data());
    Solver solver(&grid, &ca, &cb, &ham, dt, "cpu");
    for (int32_t c : calls) solver.evolve(c, imag);

    // the same evolution through HipPairKernel, as Solver::evolve drives a kernel
    State ha(&grid, 0, ar.data(), ai.data()), hb(&grid, 0, br.data(), bi.data());
    std::vector<double> ear(n), eai(n), ebr(n), ebi(n);
    exp_table(pa, grid, dt, imag, ear, eai);
    exp_table(pb, grid, dt, imag, ebr, ebi);
    double *ere[2] = {ear.data(), ebr.data()}, *eim[2] = {eai.data(), ebi.data()};
    double norm[2] = {imag ? ha.get_squared_norm() : 0., imag ? hb.get_squared_norm() : 0.};
    ITrotterKernel *k = new HipPairKernel(&grid, &ha, &hb, &ham, ere, eim, dt, norm, imag, precision);
    for (int32_t steps : calls) {
        k->rabi_coupling(0.5, dt);
        for (int i = 0; i < steps; ++i) {
            for (int c = 0; c < 2; ++c) {
                k->run_kernel_on_halo();
                if (i != steps - 1) k->start_halo_exchange();
                k->run_kernel();
                if (i != steps - 1) k->finish_halo_exchange();
                k->wait_for_completion();
            }
            k->rabi_coupling(i == steps - 1 ? 0.5 : 1., dt);
            k->normalization();
            k->cpy_first_positive_to_first_negative();
        }
    }
    std::vector<double> oar(n), oai(n), obr(n), obi(n);
    k->get_sample(grid.dim_x, 0, 0, grid.dim_x, grid.dim_y, oar.data(), oai.data(), obr.data(), obi.data());
    const std::string name = k->get_name();
    delete k;
    std::printf("{\"kernel\": \"%s\", \"rel_l2_a\": %.6e, \"rel_l2_b\": %.6e, \"solver_norm\": %.17g}\n",
                name.c_str(), rel_l2(oar, oai, car, cai), rel_l2(obr, obi, cbr, cbi), solver.get_squared_norm());
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) die("usage: dropin_check case.bin [precision]");
    const int precision = argc > 2 ? std::atoi(argv[2]) : 64;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) die("cannot open case");
    int32_t h[10];
    double par[6];
    if (std::fread(h, 4, 10, f) != 10) die("short header");
    if (h[1] == 2) return run_pair(f, h, precision);
    if (std::fread(par, 8, 6, f) != 6) die("short header");
    Lattice2D grid(h[2], par[0], h[3], par[1], h[4] != 0, h[5] != 0);
    const size_t n = (size_t)grid.dim_x * grid.dim_y;
    std::vector<double> re(n), im(n), v(n);
    if (std::fread(re.data(), 8, n, f) != n || std::fread(im.data(), 8, n, f) != n || std::fread(v.data(), 8, n, f) != n)
        die("short data");
    std::fclose(f);
    std::vector<int> calls;
    if (argc > 3) {
        for (const char *p = argv[3]; *p;) {
            calls.push_back(std::atoi(p));
            p = std::strchr(p, ',');
            if (!p) break;
            ++p;
        }
    } else {
        calls.push_back(h[6]);
    }
    const int update_at = argc > 4 ? std::atoi(argv[4]) : -1;
    const bool imag = h[7] != 0;
    const double dt = par[2];

    // reference Solver on a copy
    std::vector<double> cre(re), cim(im);
    State cstate(&grid, 0, cre.data(), cim.data());
    Potential pot(&grid, v.data());
    Hamiltonian ham(&grid, &pot, par[3], par[4], par[5]);
    Solver solver(&grid, &cstate, &ham, dt, "cpu");
    for (int c : calls) solver.evolve(c, imag);

    // the same evolution through HipKernel, as Solver::evolve drives a kernel
    State hstate(&grid, 0, re.data(), im.data());
    std::vector<double> ere(n), eim(n);
    for (int y = 0; y < grid.dim_y; ++y)
        for (int x = 0; x < grid.dim_x; ++x) {
            const double p = pot.get_value(x, y);
            const std::complex<double> e = imag ? std::exp(std::complex<double>(-dt * p, 0.))
                                                : std::exp(std::complex<double>(0., -dt * p));
            ere[(size_t)y * grid.dim_x + x] = e.real();
            eim[(size_t)y * grid.dim_x + x] = e.imag();
        }
    const double norm = imag ? hstate.get_squared_norm() : 0.;
    HipKernel *hk = new HipKernel(&grid, &hstate, &ham, ere.data(), eim.data(), dt, norm, imag, precision);
    ITrotterKernel *k = hk;
    if (tsmi_kernel_timing(hk->context(), 1) != TSMI_OK) die("kernel timing");
    int done = 0;
    std::vector<double> ore(n), oim(n);
    for (int steps : calls) {
        for (int i = 0; i < steps; ++i, ++done) {
            if (done == update_at) k->update_potential(ere.data(), eim.data(), 0);
            k->run_kernel_on_halo();
            if (i != steps - 1) k->start_halo_exchange();
            k->run_kernel();
            if (i != steps - 1) k->finish_halo_exchange();
            k->wait_for_completion();
            k->cpy_first_positive_to_first_negative();
        }
        k->get_sample(grid.dim_x, 0, 0, grid.dim_x, grid.dim_y, ore.data(), oim.data());
    }
    const double kn = k->calculate_squared_norm();
    double kt[4] = {0, 0, 0, 0};
    hk->flush();
    if (tsmi_kernel_time(hk->context(), kt) != TSMI_OK) die("kernel time");
    const std::string name = k->get_name();
    delete k;

    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        const double dr = ore[i] - cre[i], di = oim[i] - cim[i];
        num += dr * dr + di * di;
        den += cre[i] * cre[i] + cim[i] * cim[i];
    }
    std::printf("{\"kernel\": \"%s\", \"rel_l2\": %.6e, \"kernel_norm\": %.17g, \"solver_norm\": %.17g, "
                "\"fused_launches\": %d, \"single_launches\": %d}\n",
                name.c_str(), std::sqrt(num / den), kn, solver.get_squared_norm(), (int)kt[1], (int)kt[3]);
    return 0;
}
