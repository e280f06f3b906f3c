// api_timing.cpp -- times Solver::evolve through the host C++ API, in the
// caller's process (bench.py loads it with ctypes).
//
// The call pattern is the reference's own (src/solver.cpp:150-242 driven the
// way its examples drive it, e.g. examples/harmonic_osc_real.cpp): Lattice2D,
// State on caller-provided buffers, Potential from a matrix, Hamiltonian, and a
// Solver of kernel type "hip" or "hip-fp32".  The first evolve builds the kernel
// (E table on the host, uploads) and is timed on its own.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
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

// out[0] = seconds of the first evolve(warmup) (kernel construction included);
// out[1] = seconds of evolve(-steps) followed by get_squared_norm(): a soft
// update (src/solver.cpp:180-182: no copy back to the host), then the device
// reductions, which wait for every queued step; out[2] = the same with 2 steps,
// so (out[1] - out[2]) / (steps - 2) is the time per step on the device route;
// out[3] = the squared norm after the run; out[4] = seconds of evolve(steps) as
// the reference's examples call it (ending with get_sample into the State's host
// buffers, PCIe and the FP64 conversion included).  re, im, v: the whole tile
// (W*H) in the reference layout.  Returns 0, or 1 with the message in
// tsmi_api_last_error().
int tsmi_api_time_evolve(int32_t dim, double length, int32_t periodic, const double *re, const double *im,
                         const double *v, double dt, double mass, double g, int32_t precision, int32_t warmup,
                         int32_t steps, double out[5]) {
    try {
        if (steps < 3) throw std::runtime_error("tsmi_api_time_evolve: needs at least 3 steps");
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
        solver.evolve(-2);
        solver.get_squared_norm();
        out[2] = seconds_since(t);
        t = std::chrono::steady_clock::now();
        solver.evolve(-steps);
        out[3] = solver.get_squared_norm();
        out[1] = seconds_since(t);
        t = std::chrono::steady_clock::now();
        solver.evolve(steps);
        out[4] = seconds_since(t);
        return 0;
    } catch (const std::exception &e) {
        g_api_err = e.what();
        return 1;
    }
}

const char *tsmi_api_last_error(void) { return g_api_err.c_str(); }

}  // extern "C"
