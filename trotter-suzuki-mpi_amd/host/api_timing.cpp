// api_timing.cpp -- times Solver::evolve through the host C++ API, in the
// caller's process (bench.py loads it with ctypes).
//
// The call pattern is the reference's own (src/solver.cpp:150-242 driven the
// way its examples drive it, e.g. examples/harmonic_osc_real.cpp): Lattice2D,
// State on caller-provided buffers, Potential from a matrix, Hamiltonian, and a
// Solver of kernel type "hip" or "hip-fp32"; evolve(n) ends with get_sample, so
// the state is back in host memory.  The first evolve builds the kernel
// (E table on the host, uploads) and is timed on its own.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "trottersuzuki.h"

namespace {
thread_local std::string g_api_err;

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

extern "C" {

// out[0] = seconds of the first evolve(warmup) (kernel construction included),
// out[1] = seconds of evolve(steps) (its final get_sample included),
// out[2] = seconds of evolve(0) (the final get_sample alone), out[3] = the
// squared norm after the run.  re, im, v: the whole tile (W*H) in the reference
// layout.  Returns 0, or 1 with the message in tsmi_api_last_error().
int tsmi_api_time_evolve(int32_t dim, double length, int32_t periodic, const double *re, const double *im,
                         const double *v, double dt, double mass, double g, int32_t precision, int32_t warmup,
                         int32_t steps, double out[4]) {
    try {
        Lattice2D grid(dim, length, periodic != 0, periodic != 0);
        const size_t n = (size_t)grid.dim_x * (size_t)grid.dim_y;
        std::vector<double> pr(re, re + n), pi(im, im + n), pv(v, v + n);
        State state(&grid, 0, pr.data(), pi.data());
        Potential pot(&grid, pv.data());
        Hamiltonian ham(&grid, &pot, mass, g);
        Solver solver(&grid, &state, &ham, dt, precision == 32 ? "hip-fp32" : "hip");
        auto t = std::chrono::steady_clock::now();
        solver.evolve(warmup);
        out[0] = seconds_since(t);
        t = std::chrono::steady_clock::now();
        solver.evolve(steps);
        out[1] = seconds_since(t);
        t = std::chrono::steady_clock::now();
        solver.evolve(0);
        out[2] = seconds_since(t);
        out[3] = solver.get_squared_norm();
        return 0;
    } catch (const std::exception &e) {
        g_api_err = e.what();
        return 1;
    }
}

const char *tsmi_api_last_error(void) { return g_api_err.c_str(); }

}  // extern "C"
