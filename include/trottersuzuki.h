/*
 * trottersuzuki.h -- host C++ API of the MI355X Trotter-Suzuki step.
 *
 * The classes and signatures follow the reference's public header
 * (src/trottersuzuki.h:34-602) for the single-component 2D Cartesian path, so
 * that code written against the reference compiles and runs unchanged with the
 * kernel type "hip" (FP64) or "hip-fp32" (FP32 bulk, FP64 reductions).  The
 * evolution runs on the GPU through libtsmi_hip.so (include/tsmi_hip.h); this
 * layer only builds the lattice, the state and the potential factor on the host.
 *
 * Two-component systems (Hamiltonian2Component with the two-state Solver) run
 * through HipPairKernel.
 *
 * Outside this path (and rejected with std::runtime_error, as the reference's
 * my_abort does in a build without MPI): cylindrical coordinates, a rotating
 * frame (angular_velocity != 0), kernel types other than "hip"/"hip-fp32".
 */
#ifndef TSMI_TROTTERSUZUKI_H
#define TSMI_TROTTERSUZUKI_H

#include <complex>
#include <cstddef>
#include <string>

using namespace std;

/** Geometry of one process's tile (src/trottersuzuki.h:34-56). */
class Lattice {
public:
    int mpi_rank;
    int mpi_procs;
    double length_x, length_y;
    double delta_x, delta_y;
    int dim_x, dim_y;                               ///< tile size, halo included
    int global_no_halo_dim_x, global_no_halo_dim_y; ///< lattice points without halo
    int global_dim_x, global_dim_y;                 ///< lattice points with halo
    int periods[2];                                 ///< [0]: y axis periodic, [1]: x axis periodic
    string coordinate_system;
    int halo_x, halo_y;
    int start_x, start_y;
    int end_x, end_y;
    int inner_start_x, inner_start_y;
    int inner_end_x, inner_end_y;
    int mpi_coords[2], mpi_dims[2];
    virtual ~Lattice() {}
};

/** 1D chain (src/trottersuzuki.h:64-75, src/model.cpp:32-87): one row, no y
 *  halo, delta_y = 1, length_y = 0.  Cartesian only. */
class Lattice1D : public Lattice {
public:
    Lattice1D(int dim, double length, bool periodic_x_axis = false, string coordinate_system = "cartesian");
};

/** 2D lattice (src/trottersuzuki.h:83-117, src/model.cpp:103-158). */
class Lattice2D : public Lattice {
public:
    Lattice2D(int dim, double length, bool periodic_x_axis = false, bool periodic_y_axis = false,
              double angular_velocity = 0., string coordinate_system = "cartesian");
    Lattice2D(int dim_x, double length_x, int dim_y, double length_y,
              bool periodic_x_axis = false, bool periodic_y_axis = false,
              double angular_velocity = 0., string coordinate_system = "cartesian");

private:
    void init(int dim_x, double length_x, int dim_y, double length_y, bool periodic_x_axis,
              bool periodic_y_axis, double angular_velocity, string coordinate_system);
};

/** Wave function on the tile (src/trottersuzuki.h:123-175). */
class State {
public:
    double *p_real;
    double *p_imag;
    Lattice *grid;
    int angular_momentum;

    State(Lattice *grid, int angular_momentum = 0, double *p_real = 0, double *p_imag = 0);
    State(const State &obj);
    virtual ~State();
    void init_state(complex<double> (*ini_state)(double x, double y));
    void loadtxt(char *file_name);
    void imprint(complex<double> (*function)(double x, double y));
    double *get_particle_density(double *density = 0);
    double *get_phase(double *phase = 0);
    double get_expected_value(string _operator);
    double get_squared_norm(void);
    double get_mean_x(void);
    double get_mean_xx(void);
    double get_mean_y(void);
    double get_mean_yy(void);
    double get_mean_px(void);
    double get_mean_pxpx(void);
    double get_mean_py(void);
    double get_mean_pypy(void);
    double get_mean_angular_momentum(void);
    void write_to_file(string fileprefix);
    void write_particle_density(string fileprefix);
    void write_phase(string fileprefix);
    bool expected_values_updated;

protected:
    bool self_init;
    void calculate_expected_values(void);
    double mean_X, mean_XX;
    double mean_Y, mean_YY;
    double mean_Px, mean_PxPx;
    double mean_Py, mean_PyPy;
    double mean_angular_momentum;
    double norm2;
};

/** exp(i 2 pi (n_x x / L_x + n_y y / L_y)) (src/model.cpp:640-672). */
class ExponentialState : public State {
public:
    ExponentialState(Lattice1D *grid, int n_x = 1, double norm = 1, double phase = 0, double *p_real = 0,
                     double *p_imag = 0);
    ExponentialState(Lattice2D *grid, int n_x = 1, int n_y = 1, double norm = 1, double phase = 0,
                     double *p_real = 0, double *p_imag = 0);

private:
    int n_x, n_y;
    double norm, phase;
    complex<double> exp_state(double x, double y);
};

/** Gaussian (src/model.cpp:674-714). */
class GaussianState : public State {
public:
    GaussianState(Lattice1D *grid, double omega_x, double mean_x = 0, double norm = 1, double phase = 0,
                  double *p_real = 0, double *p_imag = 0);
    GaussianState(Lattice2D *grid, double omega_x, double omega_y = -1., double mean_x = 0, double mean_y = 0,
                  double norm = 1, double phase = 0, double *p_real = 0, double *p_imag = 0);

private:
    double mean_x, mean_y, omega_x, omega_y, norm, phase;
    complex<double> gauss_state(double x, double y);
};

/** Product of sines (src/model.cpp:716-748; on a Lattice1D n_y is 0, so the
 *  state is zero, as in the reference). */
class SinusoidState : public State {
public:
    SinusoidState(Lattice1D *grid, int n_x = 1, double norm = 1, double phase = 0, double *p_real = 0,
                  double *p_imag = 0);
    SinusoidState(Lattice2D *grid, int n_x = 1, int n_y = 1, double norm = 1, double phase = 0,
                  double *p_real = 0, double *p_imag = 0);

private:
    int n_x, n_y;
    double norm, phase;
    complex<double> sinusoid_state(double x, double y);
};

/** External potential (src/trottersuzuki.h:347-391, src/model.cpp:826-908). */
class Potential {
public:
    Lattice *grid;
    double *matrix;

    Potential(Lattice *grid, char *filename);
    Potential(Lattice *grid, double *external_pot = 0);
    Potential(Lattice *grid, double (*potential_function)(double x, double y));
    Potential(Lattice *grid, double (*potential_function)(double x, double y, double t), int t = 0);
    virtual ~Potential();
    virtual double get_value(int x, int y);
    bool update(double t);
    bool updated_potential_matrix;

protected:
    double current_evolution_time;
    double (*static_potential)(double x, double y);
    double (*evolving_potential)(double x, double y, double t);
    bool self_init;
    bool is_static;
};

/** 0.5 m (wx^2 (x-mx)^2 + wy^2 (y-my)^2) (src/model.cpp:910-926). */
class HarmonicPotential : public Potential {
public:
    HarmonicPotential(Lattice2D *grid, double omegax, double omegay, double mass = 1., double mean_x = 0.,
                      double mean_y = 0.);
    ~HarmonicPotential();
    double get_value(int x, int y);

private:
    double omegax, omegay, mass, mean_x, mean_y;
};

/** Single-component Hamiltonian (src/trottersuzuki.h:423-454, src/model.cpp:931-977). */
class Hamiltonian {
public:
    Potential *potential;
    double mass;
    double coupling_a;
    double LeeHuangYang_coupling_a;
    double angular_velocity;
    double rot_coord_x;
    double rot_coord_y;

    Hamiltonian(Lattice *grid, Potential *potential = 0, double mass = 1., double coupling_a = 0.,
                double LeeHuangYang_coupling_a = 0., double angular_velocity = 0., double rot_coord_x = 0,
                double rot_coord_y = 0);
    virtual ~Hamiltonian();

protected:
    bool self_init;
    Lattice *grid;
};

/** Two-component Hamiltonian (src/trottersuzuki.h:459-497, src/model.cpp:979-997). */
class Hamiltonian2Component : public Hamiltonian {
public:
    double mass_b;
    double coupling_ab;
    double coupling_b;
    double omega_r;
    double omega_i;
    Potential *potential_b;
    Hamiltonian2Component(Lattice *grid, Potential *potential = 0, Potential *potential_b = 0, double mass = 1.,
                          double mass_b = 1., double coupling_a = 0., double coupling_ab = 0.,
                          double coupling_b = 0., double omega_r = 0, double omega_i = 0,
                          double angular_velocity = 0., double rot_coord_x = 0, double rot_coord_y = 0);
    ~Hamiltonian2Component();
};

/** Kernel interface (src/trottersuzuki.h:502-520). */
class ITrotterKernel {
public:
    virtual ~ITrotterKernel() {}
    virtual void run_kernel() = 0;
    virtual void run_kernel_on_halo() = 0;
    virtual void wait_for_completion() = 0;
    virtual void get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height,
                            double *dest_real, double *dest_imag, double *dest_real2 = 0,
                            double *dest_imag2 = 0) const = 0;
    virtual void normalization() = 0;
    virtual void rabi_coupling(double var, double delta_t) = 0;
    virtual double calculate_squared_norm(bool global = true) const = 0;
    virtual bool runs_in_place() const = 0;
    virtual string get_name() const = 0;
    virtual void update_potential(double *_external_pot_real, double *_external_pot_imag, int which) = 0;
    virtual void cpy_first_positive_to_first_negative() = 0;
    virtual void start_halo_exchange() = 0;
    virtual void finish_halo_exchange() = 0;
};

/** Evolution driver (src/trottersuzuki.h:525-597, src/solver.cpp). */
class Solver {
public:
    Lattice *grid;
    State *state;
    State *state_b;
    Hamiltonian *hamiltonian;
    double current_evolution_time;

    Solver(Lattice *grid, State *state, Hamiltonian *hamiltonian, double delta_t, string kernel_type = "hip");
    Solver(Lattice *grid, State *state1, State *state2, Hamiltonian2Component *hamiltonian, double delta_t,
           string kernel_type = "hip");
    ~Solver();
    void evolve(int iterations, bool imag_time = false);
    void update_parameters();
    double get_total_energy(void);
    double get_squared_norm(size_t which = 3);
    double get_kinetic_energy(size_t which = 3);
    double get_potential_energy(size_t which = 3);
    double get_rotational_energy(size_t which = 3);
    double get_intra_species_energy(size_t which = 3);
    double get_LeeHuangYang_energy(void);
    double get_inter_species_energy(void);
    double get_rabi_energy(void);
    void set_exp_potential(double *real, int real_length, double *imag, int imag_length, int which);

private:
    bool imag_time;
    double *external_pot_real[2];
    double *external_pot_imag[2];
    double delta_t;
    double norm2[2];
    string kernel_type;
    ITrotterKernel *kernel;
    bool single_component;
    void initialize_exp_potential(double delta_t, int which = 0);
    void init_kernel();
    double energies[6];     ///< total, kinetic, potential, intra-species, LHY, squared norm
    double energies2[12];   ///< two components: as tsmi_pair_energy
    bool has_parameters_changed;
    bool energy_expected_values_updated;
    void calculate_energy_expected_values(void);
    bool is_python;
};

double const_potential(double x, double y);
void map_lattice_to_coordinate_space(Lattice *grid, int x_in, int y_in, double *x_out, double *y_out);

#endif
