// solver.cpp -- Solver of the host API: the reference's evolve loop
// (src/solver.cpp:150-242) driving HipKernel, with the energies computed on the
// device (FP64 reductions) instead of a host pass.
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "hip_kernel.h"
#include "trottersuzuki.h"

Solver::Solver(Lattice *g, State *s, Hamiltonian *h, double dt, string kt)
    : grid(g), state(s), state_b(0), hamiltonian(h), current_evolution_time(0), imag_time(false), delta_t(dt),
      norm2(0), kernel_type(kt), kernel(0), has_parameters_changed(false), energy_expected_values_updated(false),
      is_python(false) {
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    external_pot_real = new double[n];
    external_pot_imag = new double[n];
    for (int i = 0; i < 6; ++i) energies[i] = 0;
    if (kt != "hip" && kt != "hip-fp32")
        throw std::runtime_error("Unknown kernel: this library provides the kernel types \"hip\" and \"hip-fp32\"");
}

Solver::~Solver() {
    delete kernel;
    delete[] external_pot_real;
    delete[] external_pot_imag;
}

// E = exp(-i dt V) (real time) or exp(-dt V) (imaginary time) on every tile cell
// (src/solver.cpp:74-109)
void Solver::initialize_exp_potential(double dt) {
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            const double v = hamiltonian->potential->get_value(x, y);
            const complex<double> e = imag_time ? std::exp(complex<double>(-dt * v, 0.))
                                                : std::exp(complex<double>(0., -dt * v));
            external_pot_real[(size_t)y * grid->dim_x + x] = real(e);
            external_pot_imag[(size_t)y * grid->dim_x + x] = imag(e);
        }
}

void Solver::set_exp_potential(double *re, int re_len, double *im, int im_len, int which) {
    if (which != 0) throw std::runtime_error("This library evolves one component.");
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    if ((size_t)re_len > n || (size_t)im_len > n) throw std::runtime_error("set_exp_potential: array too long");
    is_python = true;
    std::memcpy(external_pot_real, re, sizeof(double) * re_len);
    std::memcpy(external_pot_imag, im, sizeof(double) * im_len);
}

// the kernel factory (src/solver.cpp:118-148) for the HIP kernel types
void Solver::init_kernel() {
    delete kernel;
    kernel = 0;
    HipKernel *k = new HipKernel(grid, state, hamiltonian, external_pot_real, external_pot_imag, delta_t, norm2,
                                 imag_time, kernel_type == "hip-fp32" ? 32 : 64);
    kernel = k;
    // raw potential for the device-side energies
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    double *v = new double[n];
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) v[(size_t)y * grid->dim_x + x] = hamiltonian->potential->get_value(x, y);
    try {
        k->set_potential(v);
    } catch (...) {
        delete[] v;
        throw;
    }
    delete[] v;
}

void Solver::evolve(int iterations, bool _imag_time) {
    if (_imag_time != imag_time || kernel == 0 || has_parameters_changed) {
        imag_time = _imag_time;
        if (imag_time) {
            initialize_exp_potential(delta_t);
            norm2 = state->get_squared_norm();
        } else if (!is_python) {
            initialize_exp_potential(delta_t);
        }
        init_kernel();
        has_parameters_changed = false;
    }
    bool soft_update = false;
    if (iterations < 0) {
        iterations = -iterations;
        soft_update = true;
    }
    for (int i = 0; i < iterations; ++i) {
        if (i > 0 && hamiltonian->potential->update(current_evolution_time)) {
            if (!is_python) initialize_exp_potential(delta_t);
            kernel->update_potential(external_pot_real, external_pot_imag, 0);
        }
        kernel->run_kernel_on_halo();
        if (i != iterations - 1) kernel->start_halo_exchange();
        kernel->run_kernel();
        if (i != iterations - 1) kernel->finish_halo_exchange();
        kernel->wait_for_completion();
        kernel->cpy_first_positive_to_first_negative();
        current_evolution_time += delta_t;
    }
    if (!soft_update) kernel->get_sample(grid->dim_x, 0, 0, grid->dim_x, grid->dim_y, state->p_real, state->p_imag);
    state->expected_values_updated = false;
    energy_expected_values_updated = false;
}

void Solver::update_parameters() { has_parameters_changed = true; }

// Solver::calculate_energy_expected_values (src/solver.cpp:244-564, one
// component) as FP64 reductions on the device over the kernel's copy of the
// state.  Before the first evolve the kernel is built from the current state.
void Solver::calculate_energy_expected_values(void) {
    if (kernel == 0) {
        if (!is_python) initialize_exp_potential(delta_t);
        norm2 = imag_time ? state->get_squared_norm() : norm2;
        init_kernel();
    }
    static_cast<HipKernel *>(kernel)->energy(energies);
    energy_expected_values_updated = true;
}

static double pick(size_t which, double value, const char *what) {
    if (which == 3 || which == 1) return value;
    if (which == 2) {
        std::cout << "The system has only one component. No input have to be given\n";
        return 0;
    }
    std::cout << "Input may be 1, 2 or 3\n";
    (void)what;
    return 0;
}

double Solver::get_total_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return energies[0];
}

double Solver::get_squared_norm(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(which, energies[5], "norm");
}

double Solver::get_kinetic_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(which, energies[1], "kinetic");
}

double Solver::get_potential_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(which, energies[2], "potential");
}

double Solver::get_rotational_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(which, 0.0, "rotational");
}

double Solver::get_intra_species_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(which, energies[3], "intra");
}

double Solver::get_LeeHuangYang_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return energies[4];
}
