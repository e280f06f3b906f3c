// solver.cpp -- Solver of the host API: the reference's evolve loop
// (src/solver.cpp:150-242) driving HipKernel (one component) or HipPairKernel
// (two components), with the energies computed on the device (FP64 reductions)
// instead of a host pass.
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "hip_kernel.h"
#include "trottersuzuki.h"

Solver::Solver(Lattice *g, State *s, Hamiltonian *h, double dt, string kt)
    : grid(g), state(s), state_b(0), hamiltonian(h), current_evolution_time(0), imag_time(false), delta_t(dt),
      kernel_type(kt), kernel(0), single_component(true), has_parameters_changed(false),
      energy_expected_values_updated(false), is_python(false) {
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    external_pot_real[0] = new double[n];
    external_pot_imag[0] = new double[n];
    external_pot_real[1] = external_pot_imag[1] = 0;
    norm2[0] = norm2[1] = 0;
    for (int i = 0; i < 6; ++i) energies[i] = 0;
    for (int i = 0; i < 12; ++i) energies2[i] = 0;
    if (kt != "hip" && kt != "hip-fp32")
        throw std::runtime_error("Unknown kernel: this library provides the kernel types \"hip\" and \"hip-fp32\"");
}

// src/solver.cpp:43-60
Solver::Solver(Lattice *g, State *s1, State *s2, Hamiltonian2Component *h, double dt, string kt)
    : grid(g), state(s1), state_b(s2), hamiltonian(h), current_evolution_time(0), imag_time(false), delta_t(dt),
      kernel_type(kt), kernel(0), single_component(false), has_parameters_changed(false),
      energy_expected_values_updated(false), is_python(false) {
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    for (int c = 0; c < 2; ++c) {
        external_pot_real[c] = new double[n];
        external_pot_imag[c] = new double[n];
    }
    norm2[0] = norm2[1] = 0;
    for (int i = 0; i < 6; ++i) energies[i] = 0;
    for (int i = 0; i < 12; ++i) energies2[i] = 0;
    if (kt != "hip" && kt != "hip-fp32")
        throw std::runtime_error("Unknown kernel: this library provides the kernel types \"hip\" and \"hip-fp32\"");
}

Solver::~Solver() {
    delete kernel;
    for (int c = 0; c < 2; ++c) {
        delete[] external_pot_real[c];
        delete[] external_pot_imag[c];
    }
}

static Potential *potential_of(Hamiltonian *h, int which) {
    return which == 0 ? h->potential : static_cast<Hamiltonian2Component *>(h)->potential_b;
}

// E = exp(-i dt V) (real time) or exp(-dt V) (imaginary time) on every tile cell
// (src/solver.cpp:74-109)
void Solver::initialize_exp_potential(double dt, int which) {
    Potential *pot = potential_of(hamiltonian, which);
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            const double v = pot->get_value(x, y);
            const complex<double> e = imag_time ? std::exp(complex<double>(-dt * v, 0.))
                                                : std::exp(complex<double>(0., -dt * v));
            external_pot_real[which][(size_t)y * grid->dim_x + x] = real(e);
            external_pot_imag[which][(size_t)y * grid->dim_x + x] = imag(e);
        }
}

void Solver::set_exp_potential(double *re, int re_len, double *im, int im_len, int which) {
    if (which != 0 && (which != 1 || single_component))
        throw std::runtime_error("set_exp_potential: no such component");
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    if ((size_t)re_len > n || (size_t)im_len > n) throw std::runtime_error("set_exp_potential: array too long");
    is_python = true;
    std::memcpy(external_pot_real[which], re, sizeof(double) * re_len);
    std::memcpy(external_pot_imag[which], im, sizeof(double) * im_len);
    // The reference's CPU kernel keeps pointers to these arrays and reads the
    // new table at its next step (src/cpukernel.cpp:252-253); the device
    // kernel holds its own copy, so the table is handed over here.
    if (kernel) kernel->update_potential(external_pot_real[which], external_pot_imag[which], which);
}

// the kernel factory (src/solver.cpp:118-148) for the HIP kernel types
void Solver::init_kernel() {
    delete kernel;
    kernel = 0;
    const int precision = kernel_type == "hip-fp32" ? 32 : 64;
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    double *v = new double[n];
    try {
        if (single_component) {
            HipKernel *k = new HipKernel(grid, state, hamiltonian, external_pot_real[0], external_pot_imag[0],
                                         delta_t, norm2[0], imag_time, precision);
            kernel = k;
            for (int y = 0; y < grid->dim_y; ++y)
                for (int x = 0; x < grid->dim_x; ++x)
                    v[(size_t)y * grid->dim_x + x] = hamiltonian->potential->get_value(x, y);
            k->set_potential(v);   // raw potential for the device-side energies
        } else {
            HipPairKernel *k = new HipPairKernel(grid, state, state_b, static_cast<Hamiltonian2Component *>(hamiltonian),
                                                 external_pot_real, external_pot_imag, delta_t, norm2, imag_time,
                                                 precision);
            kernel = k;
            for (int c = 0; c < 2; ++c) {
                Potential *pot = potential_of(hamiltonian, c);
                for (int y = 0; y < grid->dim_y; ++y)
                    for (int x = 0; x < grid->dim_x; ++x) v[(size_t)y * grid->dim_x + x] = pot->get_value(x, y);
                k->set_potential(c, v);
            }
        }
    } catch (...) {
        delete[] v;
        throw;
    }
    delete[] v;
}

// one component's run_kernel_on_halo ... wait_for_completion (src/solver.cpp:201-210)
static void component_step(ITrotterKernel *k, bool exchange) {
    k->run_kernel_on_halo();
    if (exchange) k->start_halo_exchange();
    k->run_kernel();
    if (exchange) k->finish_halo_exchange();
    k->wait_for_completion();
}

void Solver::evolve(int iterations, bool _imag_time) {
    if (_imag_time != imag_time || kernel == 0 || has_parameters_changed) {
        imag_time = _imag_time;
        if (imag_time) {
            initialize_exp_potential(delta_t, 0);
            norm2[0] = state->get_squared_norm();
            if (!single_component) {
                initialize_exp_potential(delta_t, 1);
                norm2[1] = state_b->get_squared_norm();
            }
        } else {
            if (!is_python) initialize_exp_potential(delta_t, 0);
            if (!single_component) initialize_exp_potential(delta_t, 1);
        }
        init_kernel();
        has_parameters_changed = false;
    }
    // src/solver.cpp:173-177
    double var = 0.5;
    if (!single_component && (!is_python || current_evolution_time == 0)) kernel->rabi_coupling(var, delta_t);
    var = 1.;
    bool soft_update = false;
    if (iterations < 0) {
        iterations = -iterations;
        soft_update = true;
    }
    for (int i = 0; i < iterations; ++i) {
        if (i > 0 && hamiltonian->potential->update(current_evolution_time)) {
            if (!is_python) initialize_exp_potential(delta_t, 0);
            kernel->update_potential(external_pot_real[0], external_pot_imag[0], 0);
        }
        if (!single_component && i > 0 &&
            static_cast<Hamiltonian2Component *>(hamiltonian)->potential_b->update(current_evolution_time)) {
            if (!is_python) initialize_exp_potential(delta_t, 1);
            kernel->update_potential(external_pot_real[1], external_pot_imag[1], 1);
        }
        const bool exchange = i != iterations - 1;
        component_step(kernel, exchange);
        if (!single_component) {
            component_step(kernel, exchange);
            if (i == iterations - 1) var = 0.5;
            kernel->rabi_coupling(var, delta_t);
            kernel->normalization();
        }
        kernel->cpy_first_positive_to_first_negative();
        current_evolution_time += delta_t;
    }
    if (!soft_update) {
        if (single_component)
            kernel->get_sample(grid->dim_x, 0, 0, grid->dim_x, grid->dim_y, state->p_real, state->p_imag);
        else {
            kernel->get_sample(grid->dim_x, 0, 0, grid->dim_x, grid->dim_y, state->p_real, state->p_imag,
                               state_b->p_real, state_b->p_imag);
            state_b->expected_values_updated = false;
        }
    }
    state->expected_values_updated = false;
    energy_expected_values_updated = false;
}

void Solver::update_parameters() { has_parameters_changed = true; }

// Solver::calculate_energy_expected_values (src/solver.cpp:244-564) as FP64
// reductions on the device over the kernel's copy of the state.  Before the
// first evolve the kernel is built from the current state.
void Solver::calculate_energy_expected_values(void) {
    if (kernel == 0) {
        if (!is_python) initialize_exp_potential(delta_t, 0);
        if (!single_component) initialize_exp_potential(delta_t, 1);
        if (imag_time) {
            norm2[0] = state->get_squared_norm();
            if (!single_component) norm2[1] = state_b->get_squared_norm();
        }
        init_kernel();
    }
    if (single_component)
        static_cast<HipKernel *>(kernel)->energy(energies);
    else
        static_cast<HipPairKernel *>(kernel)->energy(energies2);
    energy_expected_values_updated = true;
}

// the component selector of the getters (src/solver.cpp:572-680): 1, 2 or 3 (total)
static double pick(bool single, size_t which, double a, double b) {
    if (which == 3) return single ? a : a + b;
    if (which == 1) return a;
    if (which == 2) {
        if (single) {
            std::cout << "The system has only one component. No input have to be given\n";
            return 0;
        }
        return b;
    }
    std::cout << "Input may be 1, 2 or 3\n";
    return 0;
}

double Solver::get_total_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? energies[0] : energies2[0];
}

double Solver::get_squared_norm(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? pick(true, which, energies[5], 0) : pick(false, which, energies2[9], energies2[10]);
}

double Solver::get_kinetic_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? pick(true, which, energies[1], 0) : pick(false, which, energies2[1], energies2[2]);
}

double Solver::get_potential_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? pick(true, which, energies[2], 0) : pick(false, which, energies2[3], energies2[4]);
}

double Solver::get_rotational_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return pick(single_component, which, 0.0, 0.0);
}

double Solver::get_intra_species_energy(size_t which) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? pick(true, which, energies[3], 0) : pick(false, which, energies2[5], energies2[6]);
}

double Solver::get_LeeHuangYang_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    return single_component ? energies[4] : 0.0;   // Hamiltonian2Component has no LHY term
}

double Solver::get_inter_species_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    if (single_component) {
        std::cout << "The system has only one component\n";
        return 0;
    }
    return energies2[7];
}

double Solver::get_rabi_energy(void) {
    if (!energy_expected_values_updated) calculate_energy_expected_values();
    if (single_component) {
        std::cout << "The system has only one component\n";
        return 0;
    }
    return energies2[8];
}
