// hip_kernel.cpp -- HipKernel: ITrotterKernel on top of the C ABI of
// libtsmi_hip.so.  Builds against either trottersuzuki.h (the reference's or
// this project's); see include/hip_kernel.h.
#include "hip_kernel.h"

#include <stdexcept>
#include <string>

namespace {

void check(int status, const char *what) {
    if (status != TSMI_OK) throw std::runtime_error(std::string(what) + ": " + tsmi_last_error());
}

}  // namespace

HipKernel::HipKernel(Lattice *grid, State *state, Hamiltonian *hamiltonian, double *external_pot_real,
                     double *external_pot_imag, double delta_t, double norm, bool imag_time, int precision)
    : ctx_(nullptr), precision_(precision), width_(0), height_(0), exchange_next_(false) {
    if (grid->coordinate_system != "cartesian")
        throw std::runtime_error("The HIP kernel supports Cartesian coordinates only.");
    if (hamiltonian->angular_velocity != 0.)
        throw std::runtime_error("The HIP kernel does not work with nonzero angular velocity.");
    if (grid->mpi_procs != 1)
        throw std::runtime_error("The HIP kernel runs one tile per process on one GPU.");
    if (precision != 32 && precision != 64) throw std::runtime_error("HIP kernel precision must be 32 or 64.");
    tsmi_desc d = {};
    d.dim_x = grid->global_no_halo_dim_x;
    d.dim_y = grid->global_no_halo_dim_y;
    d.periodic_x = grid->periods[1];
    d.periodic_y = grid->periods[0];
    d.halo_x = grid->halo_x;
    d.halo_y = grid->halo_y;
    d.imag_time = imag_time ? 1 : 0;
    d.precision = precision;
    d.delta_t = delta_t;
    d.mass = hamiltonian->mass;
    d.coupling = hamiltonian->coupling_a;
    d.coupling_lhy = hamiltonian->LeeHuangYang_coupling_a;
    d.delta_x = grid->delta_x;
    d.delta_y = grid->delta_y;
    d.length_x = grid->length_x;
    d.length_y = grid->length_y;
    d.norm_target = norm;
    check(tsmi_create(&d, &ctx_), "HIP kernel");
    width_ = grid->dim_x;
    height_ = grid->dim_y;
    try {
        int32_t geo[6];
        check(tsmi_geometry(ctx_, geo), "HIP kernel");
        if (geo[0] != width_ || geo[1] != height_)
            throw std::runtime_error("HIP kernel: tile size differs from the lattice's");
        check(tsmi_set_exp_potential(ctx_, external_pot_real, imag_time ? nullptr : external_pot_imag),
              "HIP kernel");
        check(tsmi_set_state(ctx_, state->p_real, state->p_imag), "HIP kernel");
    } catch (...) {
        tsmi_destroy(ctx_);
        throw;
    }
}

HipKernel::~HipKernel() { tsmi_destroy(ctx_); }

void HipKernel::run_kernel() {
    // one step; the exchange follows unless this is the last step of an evolve
    // call (the Solver then skips start_halo_exchange, src/solver.cpp:203-209)
    check(tsmi_step(ctx_, exchange_next_ ? 1 : 0), "HIP kernel");
    exchange_next_ = false;
}

void HipKernel::run_kernel_on_halo() {}

void HipKernel::wait_for_completion() {}

void HipKernel::start_halo_exchange() { exchange_next_ = true; }

void HipKernel::get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height, double *dest_real,
                           double *dest_imag, double *dest_real2, double *dest_imag2) const {
    if (dest_real2 || dest_imag2) throw std::runtime_error("The HIP kernel evolves one component.");
    check(tsmi_get_sample(ctx_, dest_stride, x, y, width, height, dest_real, dest_imag), "HIP kernel");
}

void HipKernel::normalization() { throw std::runtime_error("The HIP kernel evolves one component."); }

void HipKernel::rabi_coupling(double, double) { throw std::runtime_error("The HIP kernel evolves one component."); }

double HipKernel::calculate_squared_norm(bool) const {
    double n = 0;
    check(tsmi_squared_norm(ctx_, &n), "HIP kernel");
    return n;
}

void HipKernel::update_potential(double *re, double *im, int which) {
    if (which != 0) throw std::runtime_error("The HIP kernel evolves one component.");
    check(tsmi_set_exp_potential(ctx_, re, im), "HIP kernel");
}

void HipKernel::evolve(int n_steps) { check(tsmi_evolve(ctx_, n_steps), "HIP kernel"); }

void HipKernel::set_potential(const double *v) { check(tsmi_set_potential(ctx_, v), "HIP kernel"); }

void HipKernel::energy(double out[6]) const { check(tsmi_energy(ctx_, out), "HIP kernel"); }

void HipKernel::expected_values(double out[10]) const { check(tsmi_expected_values(ctx_, out), "HIP kernel"); }
