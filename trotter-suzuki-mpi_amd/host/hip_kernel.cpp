// hip_kernel.cpp -- HipKernel: ITrotterKernel on top of the C ABI of
// libtsmi_hip.so.  Builds against either trottersuzuki.h (the reference's or
// this project's); see include/hip_kernel.h.
#include "hip_kernel.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace {

void check(int status, const char *what) {
    if (status != TSMI_OK) throw std::runtime_error(std::string(what) + ": " + tsmi_last_error());
}

}  // namespace

HipKernel::HipKernel(Lattice *grid, State *state, Hamiltonian *hamiltonian, double *external_pot_real,
                     double *external_pot_imag, double delta_t, double norm, bool imag_time, int precision)
    : ctx_(nullptr), precision_(precision), width_(0), height_(0), exchange_next_(false), queued_(0),
      queued_exchange_(false) {
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
    // call (the Solver then skips start_halo_exchange, src/solver.cpp:203-209).
    // Queued: a run of steps each followed by the exchange, the last one with or
    // without it, is what tsmi_run_steps issues in one call.
    if (queued_ > 0 && !queued_exchange_) flush();
    ++queued_;
    queued_exchange_ = exchange_next_;
    exchange_next_ = false;
}

void HipKernel::flush() const {
    if (queued_ == 0) return;
    const int n = queued_;
    const bool ex = queued_exchange_;
    queued_ = 0;
    queued_exchange_ = false;
    check(tsmi_run_steps(ctx_, n, ex ? 1 : 0), "HIP kernel");
}

void HipKernel::run_kernel_on_halo() {}

void HipKernel::wait_for_completion() {}

void HipKernel::start_halo_exchange() { exchange_next_ = true; }

void HipKernel::get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height, double *dest_real,
                           double *dest_imag, double *dest_real2, double *dest_imag2) const {
    if (dest_real2 || dest_imag2) throw std::runtime_error("The HIP kernel evolves one component.");
    flush();
    check(tsmi_get_sample(ctx_, dest_stride, x, y, width, height, dest_real, dest_imag), "HIP kernel");
}

void HipKernel::normalization() { throw std::runtime_error("The HIP kernel evolves one component."); }

void HipKernel::rabi_coupling(double, double) { throw std::runtime_error("The HIP kernel evolves one component."); }

double HipKernel::calculate_squared_norm(bool) const {
    flush();
    double n = 0;
    check(tsmi_squared_norm(ctx_, &n), "HIP kernel");
    return n;
}

void HipKernel::update_potential(double *re, double *im, int which) {
    if (which != 0) throw std::runtime_error("The HIP kernel evolves one component.");
    flush();
    check(tsmi_set_exp_potential(ctx_, re, im), "HIP kernel");
}

void HipKernel::evolve(int n_steps) {
    flush();
    check(tsmi_evolve(ctx_, n_steps), "HIP kernel");
}

void HipKernel::set_potential(const double *v) { check(tsmi_set_potential(ctx_, v), "HIP kernel"); }

void HipKernel::energy(double out[6]) const {
    flush();
    check(tsmi_energy(ctx_, out), "HIP kernel");
}

void HipKernel::expected_values(double out[10]) const {
    flush();
    check(tsmi_expected_values(ctx_, out), "HIP kernel");
}

// ---------------------------------------------------------------------------
// HipPairKernel
// ---------------------------------------------------------------------------

namespace {

tsmi_desc component_desc(Lattice *grid, double mass, double coupling, double delta_t, double norm, bool imag_time,
                         int precision) {
    tsmi_desc d = {};
    d.dim_x = grid->global_no_halo_dim_x;
    d.dim_y = grid->global_no_halo_dim_y;
    d.periodic_x = grid->periods[1];
    d.periodic_y = grid->periods[0];
    d.halo_x = grid->halo_x;
    d.halo_y = grid->halo_y;
    d.imag_time = imag_time ? 1 : 0;
    d.precision = precision;
    d.sweeps = 1;
    d.delta_t = delta_t;
    d.mass = mass;
    d.coupling = coupling;
    d.delta_x = grid->delta_x;
    d.delta_y = grid->delta_y;
    d.length_x = grid->length_x;
    d.length_y = grid->length_y;
    d.norm_target = norm;
    return d;
}

}  // namespace

HipPairKernel::HipPairKernel(Lattice *grid, State *state1, State *state2, Hamiltonian2Component *hamiltonian,
                             double **external_pot_real, double **external_pot_imag, double delta_t, double *norm,
                             bool imag_time, int precision)
    : pair_(nullptr), precision_(precision), width_(0), height_(0), imag_(imag_time), delta_t_(delta_t),
      state_index_(0), queued_(0), exchange_(false), rot_pending_(false), rot_var_(0) {
    if (grid->coordinate_system != "cartesian")
        throw std::runtime_error("The HIP kernel supports Cartesian coordinates only.");
    if (hamiltonian->angular_velocity != 0.)
        throw std::runtime_error("The HIP kernel does not work with nonzero angular velocity.");
    if (grid->mpi_procs != 1)
        throw std::runtime_error("The HIP kernel runs one tile per process on one GPU.");
    if (precision != 32 && precision != 64) throw std::runtime_error("HIP kernel precision must be 32 or 64.");
    tsmi_pair_desc d = {};
    // src/cpukernel.cpp:294-331: per-component mass and self coupling, then g_ab and omega
    d.a = component_desc(grid, hamiltonian->mass, hamiltonian->coupling_a, delta_t, norm[0], imag_time, precision);
    d.b = component_desc(grid, hamiltonian->mass_b, hamiltonian->coupling_b, delta_t, norm[1], imag_time, precision);
    d.coupling_ab = hamiltonian->coupling_ab;
    d.omega_r = hamiltonian->omega_r;
    d.omega_i = hamiltonian->omega_i;
    check(tsmi_pair_create(&d, &pair_), "HIP kernel");
    width_ = grid->dim_x;
    height_ = grid->dim_y;
    try {
        State *st[2] = {state1, state2};
        for (int i = 0; i < 2; ++i) {
            check(tsmi_pair_set_exp_potential(pair_, i, external_pot_real[i], imag_time ? nullptr : external_pot_imag[i]),
                  "HIP kernel");
            check(tsmi_pair_set_state(pair_, i, st[i]->p_real, st[i]->p_imag), "HIP kernel");
        }
    } catch (...) {
        tsmi_pair_destroy(pair_);
        throw;
    }
}

HipPairKernel::~HipPairKernel() { tsmi_pair_destroy(pair_); }

void HipPairKernel::run_kernel() {
    // a rotation requested before the first component's step (the half
    // rotation that opens an evolve call, src/solver.cpp:173-177) comes first
    if (state_index_ == 0 && rot_pending_) {
        check(tsmi_pair_rabi(pair_, rot_var_, 0), "HIP kernel");
        rot_pending_ = false;
    }
    if (queued_ >= 2) throw std::runtime_error("HIP kernel: a component's step was requested twice");
    ++queued_;
}

void HipPairKernel::wait_for_completion() { state_index_ = 1 - state_index_; }

void HipPairKernel::start_halo_exchange() { exchange_ = true; }

void HipPairKernel::rabi_coupling(double var, double delta_t) {
    if (delta_t != delta_t_) throw std::runtime_error("HIP kernel: rabi_coupling with a different time step");
    if (queued_ == 2) {
        // both components' steps, with the rotation in real time
        check(tsmi_pair_step(pair_, exchange_ ? 1 : 0, imag_ ? 0.0 : var), "HIP kernel");
        queued_ = 0;
        exchange_ = false;
        if (imag_) {
            rot_pending_ = true;   // after the renormalisations, at normalization()
            rot_var_ = var;
        }
        return;
    }
    if (queued_ != 0) throw std::runtime_error("HIP kernel: rabi_coupling between the two components' steps");
    if (rot_pending_) check(tsmi_pair_rabi(pair_, rot_var_, 0), "HIP kernel");
    rot_pending_ = true;
    rot_var_ = var;
}

void HipPairKernel::normalization() {
    flush();
    const double var = rot_pending_ ? rot_var_ : 0.0;
    rot_pending_ = false;
    check(tsmi_pair_rabi(pair_, var, 1), "HIP kernel");
}

void HipPairKernel::flush() const {
    if (queued_ == 1) throw std::runtime_error("HIP kernel: only one component's step was requested");
    if (queued_ == 2) {
        check(tsmi_pair_step(pair_, exchange_ ? 1 : 0, 0.0), "HIP kernel");
        queued_ = 0;
        exchange_ = false;
    }
}

void HipPairKernel::get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height,
                               double *dest_real, double *dest_imag, double *dest_real2, double *dest_imag2) const {
    flush();
    if (rot_pending_) {
        check(tsmi_pair_rabi(pair_, rot_var_, 0), "HIP kernel");
        rot_pending_ = false;
    }
    if (x + width > (size_t)width_ || y + height > (size_t)height_ || dest_stride < width)
        throw std::runtime_error("HIP kernel: sample window outside the tile");
    const size_t n = (size_t)width_ * (size_t)height_;
    std::vector<double> re(n), im(n);
    double *dr[2] = {dest_real, dest_real2}, *di[2] = {dest_imag, dest_imag2};
    for (int c = 0; c < 2; ++c) {
        if (!dr[c] || !di[c]) continue;
        check(tsmi_pair_get_state(pair_, c, re.data(), im.data()), "HIP kernel");
        for (size_t j = 0; j < height; ++j)
            for (size_t i = 0; i < width; ++i) {
                dr[c][j * dest_stride + i] = re[(y + j) * (size_t)width_ + x + i];
                di[c][j * dest_stride + i] = im[(y + j) * (size_t)width_ + x + i];
            }
    }
}

double HipPairKernel::calculate_squared_norm(bool) const {
    double e[12];
    energy(e);
    return e[9 + state_index_];
}

void HipPairKernel::update_potential(double *re, double *im, int which) {
    if (which != 0 && which != 1) throw std::runtime_error("HIP kernel: component 0 or 1");
    flush();
    check(tsmi_pair_set_exp_potential(pair_, which, re, imag_ ? nullptr : im), "HIP kernel");
}

void HipPairKernel::set_potential(int which, const double *v) {
    check(tsmi_pair_set_potential(pair_, which, v), "HIP kernel");
}

void HipPairKernel::energy(double out[12]) const {
    flush();
    if (rot_pending_) {
        check(tsmi_pair_rabi(pair_, rot_var_, 0), "HIP kernel");
        rot_pending_ = false;
    }
    check(tsmi_pair_energy(pair_, out), "HIP kernel");
}
