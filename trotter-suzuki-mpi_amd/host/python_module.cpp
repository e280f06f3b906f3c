// python_module.cpp -- the compiled part of the `trottersuzuki` Python package:
// the host C++ API (include/trottersuzuki.h) as Python classes with the names,
// constructors, attributes and methods the reference's SWIG module exposes
// (src/Python/trottersuzuki/trottersuzuki.i:37-333), numpy arrays where that
// interface maps arrays (init_state_matrix, imprint_matrix,
// init_potential_matrix, get_particle_density, get_phase, set_exp_potential).
// The Python-level additions of the reference (init_state, init_potential,
// the time-dependent evolve loop, ...) live in trottersuzuki/classes.py.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "trottersuzuki.h"

namespace py = pybind11;

namespace {

using Arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

// A (dim_y, dim_x) array over the whole tile, or its flat equivalent.
const double *tile_data(const Arr &a, const Lattice *g, const char *what) {
    const size_t n = (size_t)g->dim_x * g->dim_y;
    if ((size_t)a.size() != n)
        throw std::runtime_error(std::string(what) + ": expected " + std::to_string(g->dim_y) + " x " +
                                 std::to_string(g->dim_x) + " values (the whole tile)");
    return a.data();
}

// A (rows, cols) numpy array that owns a copy of `src`.
py::array_t<double> owned(const double *src, int rows, int cols) {
    py::array_t<double> out({rows, cols});
    std::memcpy(out.mutable_data(), src, sizeof(double) * (size_t)rows * cols);
    return out;
}

}  // namespace

PYBIND11_MODULE(_trottersuzuki, m) {
    m.doc() = "Host API of the MI355X Trotter-Suzuki step (HIP kernels through libtsmi_hip.so)";

    py::class_<Lattice>(m, "Lattice")
        .def_readwrite("length_x", &Lattice::length_x)
        .def_readwrite("length_y", &Lattice::length_y)
        .def_readwrite("delta_x", &Lattice::delta_x)
        .def_readwrite("delta_y", &Lattice::delta_y)
        .def_readwrite("dim_x", &Lattice::dim_x)
        .def_readwrite("dim_y", &Lattice::dim_y)
        .def_readwrite("global_no_halo_dim_x", &Lattice::global_no_halo_dim_x)
        .def_readwrite("global_no_halo_dim_y", &Lattice::global_no_halo_dim_y)
        .def_readwrite("start_x", &Lattice::start_x)
        .def_readwrite("start_y", &Lattice::start_y)
        .def_readwrite("inner_start_x", &Lattice::inner_start_x)
        .def_readwrite("inner_start_y", &Lattice::inner_start_y)
        .def_readwrite("inner_end_x", &Lattice::inner_end_x)
        .def_readwrite("inner_end_y", &Lattice::inner_end_y)
        .def_readwrite("halo_x", &Lattice::halo_x)
        .def_readwrite("halo_y", &Lattice::halo_y)
        .def_readwrite("coordinate_system", &Lattice::coordinate_system)
        .def_property_readonly("periodic_x_axis", [](const Lattice &g) { return g.periods[1] != 0; })
        .def_property_readonly("periodic_y_axis", [](const Lattice &g) { return g.periods[0] != 0; });

    py::class_<Lattice1D, Lattice>(m, "Lattice1D")
        .def(py::init<int, double, bool, std::string>(), py::arg("dim"), py::arg("length"),
             py::arg("periodic_x_axis") = false, py::arg("coordinate_system") = "cartesian");

    py::class_<Lattice2D, Lattice>(m, "Lattice2D")
        .def(py::init<int, double, int, double, bool, bool, double, std::string>(), py::arg("dim_x"),
             py::arg("length_x"), py::arg("dim_y"), py::arg("length_y"), py::arg("periodic_x_axis") = false,
             py::arg("periodic_y_axis") = false, py::arg("angular_velocity") = 0.,
             py::arg("coordinate_system") = "cartesian");

    py::class_<State>(m, "State")
        .def(py::init([](Lattice *g, int am) { return new State(g, am); }), py::arg("grid"),
             py::arg("angular_momentum") = 0, py::keep_alive<1, 2>())
        .def(py::init([](const State &o) { return new State(o); }), py::arg("obj"))
        .def_readonly("grid", &State::grid, py::return_value_policy::reference)
        .def_readwrite("expected_values_updated", &State::expected_values_updated)
        .def("init_state_matrix",
             [](State &s, const Arr &re, const Arr &im) {
                 const size_t n = (size_t)s.grid->dim_x * s.grid->dim_y;
                 std::memcpy(s.p_real, tile_data(re, s.grid, "init_state_matrix"), n * sizeof(double));
                 std::memcpy(s.p_imag, tile_data(im, s.grid, "init_state_matrix"), n * sizeof(double));
                 s.expected_values_updated = false;
             },
             py::arg("state_real"), py::arg("state_imag"))
        .def("imprint_matrix",
             [](State &s, const Arr &re, const Arr &im) {
                 const double *a = tile_data(re, s.grid, "imprint_matrix");
                 const double *b = tile_data(im, s.grid, "imprint_matrix");
                 const size_t n = (size_t)s.grid->dim_x * s.grid->dim_y;
                 for (size_t k = 0; k < n; ++k) {
                     const double r = s.p_real[k];
                     s.p_real[k] = r * a[k] - s.p_imag[k] * b[k];
                     s.p_imag[k] = r * b[k] + s.p_imag[k] * a[k];
                 }
                 s.expected_values_updated = false;
             },
             py::arg("state_real"), py::arg("state_imag"))
        .def("loadtxt", [](State &s, std::string f) { s.loadtxt(&f[0]); }, py::arg("file_name"))
        .def("get_particle_density",
             [](State &s) {
                 std::unique_ptr<double[]> d(s.get_particle_density());
                 return owned(d.get(), s.grid->inner_end_y - s.grid->inner_start_y,
                              s.grid->inner_end_x - s.grid->inner_start_x);
             })
        .def("get_phase",
             [](State &s) {
                 std::unique_ptr<double[]> d(s.get_phase());
                 return owned(d.get(), s.grid->inner_end_y - s.grid->inner_start_y,
                              s.grid->inner_end_x - s.grid->inner_start_x);
             })
        .def("get_expected_value", &State::get_expected_value, py::arg("_operator"))
        .def("get_squared_norm", &State::get_squared_norm)
        .def("get_mean_x", &State::get_mean_x)
        .def("get_mean_xx", &State::get_mean_xx)
        .def("get_mean_y", &State::get_mean_y)
        .def("get_mean_yy", &State::get_mean_yy)
        .def("get_mean_px", &State::get_mean_px)
        .def("get_mean_pxpx", &State::get_mean_pxpx)
        .def("get_mean_py", &State::get_mean_py)
        .def("get_mean_pypy", &State::get_mean_pypy)
        .def("get_mean_angular_momentum", &State::get_mean_angular_momentum)
        .def("write_to_file", &State::write_to_file, py::arg("fileprefix"))
        .def("write_particle_density", &State::write_particle_density, py::arg("fileprefix"))
        .def("write_phase", &State::write_phase, py::arg("fileprefix"))
        // the whole tile (halo included) as (dim_y, dim_x) copies, for inspection
        .def("_tile",
             [](State &s) {
                 return py::make_tuple(owned(s.p_real, s.grid->dim_y, s.grid->dim_x),
                                       owned(s.p_imag, s.grid->dim_y, s.grid->dim_x));
             });

    py::class_<ExponentialState, State>(m, "ExponentialState")
        .def(py::init([](Lattice1D *g, int nx, double norm, double phase) {
                 return new ExponentialState(g, nx, norm, phase);
             }),
             py::arg("grid"), py::arg("n_x") = 1, py::arg("norm") = 1., py::arg("phase") = 0.,
             py::keep_alive<1, 2>())
        .def(py::init([](Lattice2D *g, int nx, int ny, double norm, double phase) {
                 return new ExponentialState(g, nx, ny, norm, phase);
             }),
             py::arg("grid"), py::arg("n_x") = 1, py::arg("n_y") = 1, py::arg("norm") = 1., py::arg("phase") = 0.,
             py::keep_alive<1, 2>());

    py::class_<GaussianState, State>(m, "GaussianState")
        .def(py::init([](Lattice1D *g, double ox, double mx, double norm, double phase) {
                 return new GaussianState(g, ox, mx, norm, phase);
             }),
             py::arg("grid"), py::arg("omega_x"), py::arg("mean_x") = 0., py::arg("norm") = 1.,
             py::arg("phase") = 0., py::keep_alive<1, 2>())
        .def(py::init([](Lattice2D *g, double ox, double oy, double mx, double my, double norm, double phase) {
                 return new GaussianState(g, ox, oy, mx, my, norm, phase);
             }),
             py::arg("grid"), py::arg("omega_x"), py::arg("omega_y") = -1., py::arg("mean_x") = 0.,
             py::arg("mean_y") = 0., py::arg("norm") = 1., py::arg("phase") = 0., py::keep_alive<1, 2>());

    py::class_<SinusoidState, State>(m, "SinusoidState")
        .def(py::init([](Lattice1D *g, int nx, double norm, double phase) {
                 return new SinusoidState(g, nx, norm, phase);
             }),
             py::arg("grid"), py::arg("n_x") = 1, py::arg("norm") = 1., py::arg("phase") = 0.,
             py::keep_alive<1, 2>())
        .def(py::init([](Lattice2D *g, int nx, int ny, double norm, double phase) {
                 return new SinusoidState(g, nx, ny, norm, phase);
             }),
             py::arg("grid"), py::arg("n_x") = 1, py::arg("n_y") = 1, py::arg("norm") = 1., py::arg("phase") = 0.,
             py::keep_alive<1, 2>());

    py::class_<Potential>(m, "Potential")
        .def(py::init([](Lattice *g) { return new Potential(g, (double *)nullptr); }), py::arg("grid"),
             py::keep_alive<1, 2>())
        .def(py::init([](Lattice *g, std::string f) { return new Potential(g, &f[0]); }), py::arg("grid"),
             py::arg("filename"), py::keep_alive<1, 2>())
        .def_readonly("grid", &Potential::grid, py::return_value_policy::reference)
        .def_readwrite("updated_potential_matrix", &Potential::updated_potential_matrix)
        .def("init_potential_matrix",
             [](Potential &p, const Arr &v) {
                 const size_t n = (size_t)p.grid->dim_x * p.grid->dim_y;
                 std::memcpy(p.matrix, tile_data(v, p.grid, "init_potential_matrix"), n * sizeof(double));
             },
             py::arg("_potential"))
        .def("get_value", &Potential::get_value, py::arg("x"), py::arg("y"))
        .def("update", &Potential::update, py::arg("t"));

    py::class_<HarmonicPotential, Potential>(m, "HarmonicPotential")
        .def(py::init<Lattice2D *, double, double, double, double, double>(), py::arg("grid"), py::arg("omegax"),
             py::arg("omegay"), py::arg("mass") = 1., py::arg("mean_x") = 0., py::arg("mean_y") = 0.,
             py::keep_alive<1, 2>());

    py::class_<Hamiltonian>(m, "Hamiltonian")
        .def(py::init<Lattice *, Potential *, double, double, double, double, double, double>(), py::arg("grid"),
             py::arg("potential") = nullptr, py::arg("mass") = 1., py::arg("coupling_a") = 0.,
             py::arg("LeeHuangYang_coupling_a") = 0., py::arg("angular_velocity") = 0.,
             py::arg("rot_coord_x") = 0., py::arg("rot_coord_y") = 0., py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>())
        .def_readonly("potential", &Hamiltonian::potential, py::return_value_policy::reference)
        .def_readwrite("mass", &Hamiltonian::mass)
        .def_readwrite("coupling_a", &Hamiltonian::coupling_a)
        .def_readwrite("LeeHuangYang_coupling_a", &Hamiltonian::LeeHuangYang_coupling_a)
        .def_readwrite("angular_velocity", &Hamiltonian::angular_velocity)
        .def_readwrite("rot_coord_x", &Hamiltonian::rot_coord_x)
        .def_readwrite("rot_coord_y", &Hamiltonian::rot_coord_y);

    py::class_<Hamiltonian2Component, Hamiltonian>(m, "Hamiltonian2Component")
        .def(py::init<Lattice *, Potential *, Potential *, double, double, double, double, double, double, double,
                      double, double, double>(),
             py::arg("grid"), py::arg("potential") = nullptr, py::arg("potential_b") = nullptr,
             py::arg("mass") = 1., py::arg("mass_b") = 1., py::arg("coupling_a") = 0., py::arg("coupling_ab") = 0.,
             py::arg("coupling_b") = 0., py::arg("omega_r") = 0., py::arg("omega_i") = 0.,
             py::arg("angular_velocity") = 0., py::arg("rot_coord_x") = 0., py::arg("rot_coord_y") = 0.,
             py::keep_alive<1, 2>(), py::keep_alive<1, 3>(), py::keep_alive<1, 4>())
        .def_readonly("potential_b", &Hamiltonian2Component::potential_b, py::return_value_policy::reference)
        .def_readwrite("mass_b", &Hamiltonian2Component::mass_b)
        .def_readwrite("coupling_ab", &Hamiltonian2Component::coupling_ab)
        .def_readwrite("coupling_b", &Hamiltonian2Component::coupling_b)
        .def_readwrite("omega_r", &Hamiltonian2Component::omega_r)
        .def_readwrite("omega_i", &Hamiltonian2Component::omega_i);

    py::class_<Solver>(m, "Solver")
        .def(py::init<Lattice *, State *, Hamiltonian *, double, std::string>(), py::arg("grid"), py::arg("state"),
             py::arg("hamiltonian"), py::arg("delta_t"), py::arg("kernel_type") = "hip", py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>(), py::keep_alive<1, 4>())
        .def(py::init<Lattice *, State *, State *, Hamiltonian2Component *, double, std::string>(), py::arg("grid"),
             py::arg("state1"), py::arg("state2"), py::arg("hamiltonian"), py::arg("delta_t"),
             py::arg("kernel_type") = "hip", py::keep_alive<1, 2>(), py::keep_alive<1, 3>(),
             py::keep_alive<1, 4>(), py::keep_alive<1, 5>())
        .def_readonly("grid", &Solver::grid, py::return_value_policy::reference)
        .def_readonly("state", &Solver::state, py::return_value_policy::reference)
        .def_readonly("state_b", &Solver::state_b, py::return_value_policy::reference)
        .def_readonly("hamiltonian", &Solver::hamiltonian, py::return_value_policy::reference)
        .def_readwrite("current_evolution_time", &Solver::current_evolution_time)
        .def("evolve", &Solver::evolve, py::arg("iterations"), py::arg("imag_time") = false,
             py::call_guard<py::gil_scoped_release>())
        .def("update_parameters", &Solver::update_parameters)
        .def("get_total_energy", &Solver::get_total_energy)
        .def("get_squared_norm", &Solver::get_squared_norm, py::arg("which") = 3)
        .def("get_kinetic_energy", &Solver::get_kinetic_energy, py::arg("which") = 3)
        .def("get_potential_energy", &Solver::get_potential_energy, py::arg("which") = 3)
        .def("get_rotational_energy", &Solver::get_rotational_energy, py::arg("which") = 3)
        .def("get_intra_species_energy", &Solver::get_intra_species_energy, py::arg("which") = 3)
        .def("get_LeeHuangYang_energy", &Solver::get_LeeHuangYang_energy)
        .def("get_inter_species_energy", &Solver::get_inter_species_energy)
        .def("get_rabi_energy", &Solver::get_rabi_energy)
        .def("set_exp_potential",
             [](Solver &s, const Arr &re, const Arr &im, int which) {
                 const size_t n = (size_t)s.grid->dim_x * s.grid->dim_y;
                 if ((size_t)re.size() > n || (size_t)im.size() > n)
                     throw std::runtime_error("set_exp_potential: array longer than the tile");
                 s.set_exp_potential(const_cast<double *>(re.data()), (int)re.size(), const_cast<double *>(im.data()),
                                     (int)im.size(), which);
             },
             py::arg("exp_pot_real"), py::arg("exp_pot_imag"), py::arg("which"));
}
