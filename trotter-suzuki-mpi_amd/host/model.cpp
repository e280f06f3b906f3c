// model.cpp -- lattice, states, potentials and Hamiltonian of the host API
// (include/trottersuzuki.h).  Semantics follow the reference's model.cpp and
// common.cpp for one process; cited per function.
#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "trottersuzuki.h"

namespace {

const int kHalo = 4;   // one step's dependency cone (src/model.cpp:140-141, no rotation)

// Tile bounds of one axis for a single process (src/common.cpp:83-102 with one rank).
void borders(int length, int halo, bool periodic, int *start, int *end, int *inner_start, int *inner_end) {
    *inner_start = 0;
    *start = periodic ? -halo : 0;
    *end = periodic ? length + halo : length;
    *inner_end = length;
}

void write_matrix(const std::string &name, const double *m, int stride, int w, int h) {
    std::ofstream out(name.c_str(), std::ios::out | std::ios::trunc);
    for (int i = 0; i < h; ++i) {
        for (int j = 0; j < w; ++j) out << m[i * stride + j] << " ";
        out << std::endl;
    }
}

}  // namespace

double const_potential(double, double) { return 0.; }

// src/common.cpp:45-65 (Cartesian)
void map_lattice_to_coordinate_space(Lattice *grid, int x_in, int y_in, double *x_out, double *y_out) {
    double idy = grid->start_y * grid->delta_y + 0.5 * grid->delta_y + y_in * grid->delta_y;
    double idx = grid->start_x * grid->delta_x + 0.5 * grid->delta_x + x_in * grid->delta_x;
    const double x_c = grid->global_no_halo_dim_x * grid->delta_x * 0.5;
    const double y_c = grid->global_no_halo_dim_y * grid->delta_y * 0.5;
    if (idx - x_c < -grid->length_x * 0.5) idx += grid->length_x;
    if (idx - x_c > grid->length_x * 0.5) idx -= grid->length_x;
    if (idy - y_c < -grid->length_y * 0.5) idy += grid->length_y;
    if (idy - y_c > grid->length_y * 0.5) idy -= grid->length_y;
    *x_out = idx - x_c;
    *y_out = idy - y_c;
}

// ---------------------------------------------------------------------------
// Lattice1D (src/model.cpp:32-87): one row, no y halo, delta_y = 1, length_y = 0
// ---------------------------------------------------------------------------
Lattice1D::Lattice1D(int dim, double length, bool px, string cs) {
    if (cs != "cartesian" && cs != "cylindrical")
        throw std::runtime_error("The coordinate system you have chosen is not implemented.");
    if (cs == "cylindrical") throw std::runtime_error("Cylindrical coordinates are outside the HIP path.");
    if (dim < 1) throw std::runtime_error("Lattice dimensions must be positive.");
    coordinate_system = cs;
    length_x = length;
    length_y = 0;
    delta_x = length_x / double(dim);
    delta_y = 1.0;
    periods[0] = 0;
    periods[1] = (int)px;
    mpi_procs = 1;
    mpi_rank = 0;
    mpi_dims[0] = mpi_dims[1] = 1;
    mpi_coords[0] = mpi_coords[1] = 0;
    halo_x = kHalo;
    halo_y = 0;
    global_dim_x = dim + periods[1] * 2 * halo_x;
    global_dim_y = 1;
    global_no_halo_dim_x = dim;
    global_no_halo_dim_y = 1;
    borders(dim, halo_x, px, &start_x, &end_x, &inner_start_x, &inner_end_x);
    dim_x = end_x - start_x;
    start_y = 0;
    end_y = 1;
    inner_start_y = 0;
    inner_end_y = 1;
    dim_y = 1;
}

// ---------------------------------------------------------------------------
// Lattice2D (src/model.cpp:89-158)
// ---------------------------------------------------------------------------
Lattice2D::Lattice2D(int dim, double length, bool px, bool py, double omega, string cs) {
    init(dim, length, dim, length, px, py, omega, cs);
}

Lattice2D::Lattice2D(int dim_x_, double length_x_, int dim_y_, double length_y_, bool px, bool py, double omega,
                     string cs) {
    init(dim_x_, length_x_, dim_y_, length_y_, px, py, omega, cs);
}

void Lattice2D::init(int nx, double lx, int ny, double ly, bool px, bool py, double omega, string cs) {
    if (cs != "cartesian" && cs != "cylindrical")
        throw std::runtime_error("The coordinate system you have chosen is not implemented.");
    if (cs == "cylindrical") throw std::runtime_error("Cylindrical coordinates are outside the HIP path.");
    if (omega != 0.) throw std::runtime_error("A rotating frame (angular velocity != 0) is outside the HIP path.");
    if (nx < 1 || ny < 1) throw std::runtime_error("Lattice dimensions must be positive.");
    length_x = lx;
    length_y = ly;
    delta_x = length_x / double(nx);
    delta_y = length_y / double(ny);
    coordinate_system = cs;
    periods[0] = (int)py;
    periods[1] = (int)px;
    mpi_procs = 1;
    mpi_rank = 0;
    mpi_dims[0] = mpi_dims[1] = 1;
    mpi_coords[0] = mpi_coords[1] = 0;
    halo_x = kHalo;
    halo_y = kHalo;
    global_dim_x = nx + periods[1] * 2 * halo_x;
    global_dim_y = ny + periods[0] * 2 * halo_y;
    global_no_halo_dim_x = nx;
    global_no_halo_dim_y = ny;
    borders(nx, halo_x, px, &start_x, &end_x, &inner_start_x, &inner_end_x);
    borders(ny, halo_y, py, &start_y, &end_y, &inner_start_y, &inner_end_y);
    dim_x = end_x - start_x;
    dim_y = end_y - start_y;
}

// ---------------------------------------------------------------------------
// State (src/model.cpp:160-638)
// ---------------------------------------------------------------------------
State::State(Lattice *g, int am, double *re, double *im) : grid(g), angular_momentum(am) {
    expected_values_updated = false;
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    self_init = re == 0;
    p_real = re ? re : new double[n]();
    p_imag = im ? im : new double[n]();
    mean_X = mean_XX = mean_Y = mean_YY = mean_Px = mean_PxPx = mean_Py = mean_PyPy = 0;
    mean_angular_momentum = norm2 = 0;
}

State::State(const State &o)
    : grid(o.grid), angular_momentum(o.angular_momentum), expected_values_updated(o.expected_values_updated),
      self_init(true), mean_X(o.mean_X), mean_XX(o.mean_XX), mean_Y(o.mean_Y), mean_YY(o.mean_YY), mean_Px(o.mean_Px),
      mean_PxPx(o.mean_PxPx), mean_Py(o.mean_Py), mean_PyPy(o.mean_PyPy),
      mean_angular_momentum(o.mean_angular_momentum), norm2(o.norm2) {
    const size_t n = (size_t)grid->dim_x * grid->dim_y;
    p_real = new double[n];
    p_imag = new double[n];
    for (size_t i = 0; i < n; ++i) {
        p_real[i] = o.p_real[i];
        p_imag[i] = o.p_imag[i];
    }
}

State::~State() {
    if (self_init) {
        delete[] p_real;
        delete[] p_imag;
    }
}

void State::init_state(complex<double> (*f)(double, double)) {
    double xr, yr;
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
            const complex<double> v = f(xr, yr);
            p_real[y * grid->dim_x + x] = real(v);
            p_imag[y * grid->dim_x + x] = imag(v);
        }
    expected_values_updated = false;
}

void State::imprint(complex<double> (*f)(double, double)) {
    double xr, yr;
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
            const complex<double> v = f(xr, yr);
            const size_t k = (size_t)y * grid->dim_x + x;
            const double r = p_real[k];
            p_real[k] = r * real(v) - p_imag[k] * imag(v);
            p_imag[k] = r * imag(v) + p_imag[k] * real(v);
        }
    expected_values_updated = false;
}

// Reads the inner lattice as whitespace-separated complex numbers "(re,im)" and
// fills the periodic halo from the opposite edge (src/model.cpp:254-322, one process).
void State::loadtxt(char *file_name) {
    std::ifstream in(file_name);
    if (!in) throw std::runtime_error(std::string("cannot open ") + file_name);
    const int nx = grid->global_no_halo_dim_x, ny = grid->global_no_halo_dim_y;
    const int hx = grid->periods[1] ? grid->halo_x : 0, hy = grid->periods[0] ? grid->halo_y : 0;
    const int W = grid->dim_x, H = grid->dim_y;
    complex<double> v;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            in >> v;
            // every tile cell that holds lattice point (j, i), halo copies included
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dy && !hy) || (dx && !hx)) continue;
                    const int ty = i + hy + dy * ny, tx = j + hx + dx * nx;
                    if (ty < 0 || ty >= H || tx < 0 || tx >= W) continue;
                    p_real[(size_t)ty * W + tx] = real(v);
                    p_imag[(size_t)ty * W + tx] = imag(v);
                }
        }
    expected_values_updated = false;
}

double *State::get_particle_density(double *d) {
    const int w = grid->inner_end_x - grid->inner_start_x, h = grid->inner_end_y - grid->inner_start_y;
    if (!d) d = new double[(size_t)w * h];
    for (int j = 0, y = grid->inner_start_y - grid->start_y; j < h; ++j, ++y)
        for (int i = 0, x = grid->inner_start_x - grid->start_x; i < w; ++i, ++x) {
            const size_t k = (size_t)y * grid->dim_x + x;
            d[(size_t)j * w + i] = p_real[k] * p_real[k] + p_imag[k] * p_imag[k];
        }
    return d;
}

double *State::get_phase(double *ph) {
    const int w = grid->inner_end_x - grid->inner_start_x, h = grid->inner_end_y - grid->inner_start_y;
    if (!ph) ph = new double[(size_t)w * h];
    for (int j = 0, y = grid->inner_start_y - grid->start_y; j < h; ++j, ++y)
        for (int i = 0, x = grid->inner_start_x - grid->start_x; i < w; ++i, ++x) {
            const size_t k = (size_t)y * grid->dim_x + x;
            const double n = std::sqrt(p_real[k] * p_real[k] + p_imag[k] * p_imag[k]);
            ph[(size_t)j * w + i] = n == 0 ? 0 : std::acos(p_real[k] / n) * ((p_imag[k] >= 0) - (p_imag[k] < 0));
        }
    return ph;
}

// src/model.cpp:385-536, one process
void State::calculate_expected_values(void) {
    const int W = grid->dim_x;
    const int x0 = grid->inner_start_x - grid->start_x, x1 = grid->inner_end_x - grid->start_x;
    const int y0 = grid->inner_start_y - grid->start_y, y1 = grid->inner_end_y - grid->start_y;
    const int mx = grid->periods[1] ? 0 : 2, my = grid->periods[0] ? 0 : 2;
    const double c1 = -1. / 12., c2 = 4. / 3., c3 = -2.5;
    const double d1 = 1. / 6., d2 = -1., d3 = 0.5, d4 = 1. / 3.;
    double sn = 0, sx = 0, sxx = 0, sy = 0, syy = 0, spx = 0, spy = 0, spxpx = 0, spypy = 0, slz = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            double X, Y;
            map_lattice_to_coordinate_space(grid, x, y, &X, &Y);
            const size_t k = (size_t)y * W + x;
            const double pr = p_real[k], pi = p_imag[k], n = pr * pr + pi * pi;
            sn += n; sx += n * X; sy += n * Y; sxx += n * X * X; syy += n * Y * Y;
            if (!(y >= y0 + my && y < y1 - my && x >= x0 + mx && x < x1 - mx)) continue;
            auto R = [&](long o) { return p_real[(long)k + o]; };
            auto I = [&](long o) { return p_imag[(long)k + o]; };
            const double dxr = d4 * R(1) + d3 * pr + d2 * R(-1) + d1 * R(-2);
            const double dxi = d4 * I(1) + d3 * pi + d2 * I(-1) + d1 * I(-2);
            const double dyr = d4 * R(-W) + d3 * pr + d2 * R(W) + d1 * R(2 * W);
            const double dyi = d4 * I(-W) + d3 * pi + d2 * I(W) + d1 * I(2 * W);
            spx += pr * dxi - pi * dxr;
            spy += pr * dyi - pi * dyr;
            spxpx += pr * (c1 * R(2) + c2 * R(1) + c2 * R(-1) + c1 * R(-2) + c3 * pr) +
                     pi * (c1 * I(2) + c2 * I(1) + c2 * I(-1) + c1 * I(-2) + c3 * pi);
            spypy += pr * (c1 * R(2 * W) + c2 * R(W) + c2 * R(-W) + c1 * R(-2 * W) + c3 * pr) +
                     pi * (c1 * I(2 * W) + c2 * I(W) + c2 * I(-W) + c1 * I(-2 * W) + c3 * pi);
            const double zr = Y / grid->delta_x * dxr + X / grid->delta_y * dyr;
            const double zi = Y / grid->delta_x * dxi + X / grid->delta_y * dyi;
            slz += pr * zi - pi * zr;
        }
    const double ppx = -1. / grid->delta_x, ppy = 1. / grid->delta_y;
    mean_X = sx / sn;
    mean_Y = sy / sn;
    mean_XX = sxx / sn;
    mean_YY = syy / sn;
    mean_Px = -spx * ppx / sn;
    mean_Py = -spy * ppy / sn;
    mean_PxPx = -spxpx * ppx * ppx / sn;
    mean_PyPy = -spypy * ppy * ppy / sn;
    mean_angular_momentum = slz / sn;
    norm2 = sn * (grid->delta_x * grid->delta_y);
    expected_values_updated = true;
}

double State::get_expected_value(string op) {
    if (!expected_values_updated) calculate_expected_values();
    if (op == "L_z") return mean_angular_momentum;
    if (op == "X") return mean_X;
    if (op == "X^2") return mean_XX;
    if (op == "Y") return mean_Y;
    if (op == "Y^2") return mean_YY;
    if (op == "P_x") return mean_Px;
    if (op == "P_x^2") return mean_PxPx;
    if (op == "P_y") return mean_Py;
    if (op == "P_y^2") return mean_PyPy;
    std::cerr << "The expected value of the operator " << op << " is not calculated.\n"
              << "Available operators are: L_z, X, X^2, Y, Y^2, P_x, P_x^2, P_y, P_y^2.\n";
    return std::numeric_limits<double>::quiet_NaN();
}

#define TSMI_STATE_GETTER(fn, field)                          \
    double State::fn(void) {                                  \
        if (!expected_values_updated) calculate_expected_values(); \
        return field;                                         \
    }
TSMI_STATE_GETTER(get_squared_norm, norm2)
TSMI_STATE_GETTER(get_mean_x, mean_X)
TSMI_STATE_GETTER(get_mean_xx, mean_XX)
TSMI_STATE_GETTER(get_mean_y, mean_Y)
TSMI_STATE_GETTER(get_mean_yy, mean_YY)
TSMI_STATE_GETTER(get_mean_px, mean_Px)
TSMI_STATE_GETTER(get_mean_pxpx, mean_PxPx)
TSMI_STATE_GETTER(get_mean_py, mean_Py)
TSMI_STATE_GETTER(get_mean_pypy, mean_PyPy)
TSMI_STATE_GETTER(get_mean_angular_momentum, mean_angular_momentum)
#undef TSMI_STATE_GETTER

// Text output of the inner lattice, "(re,im) " per entry and one row per line,
// in the stream's default format (src/common.cpp:161-170 via stamp, one process).
void State::write_to_file(string prefix) {
    std::ofstream out(prefix.c_str(), std::ios::out | std::ios::trunc);
    const int x0 = grid->inner_start_x - grid->start_x, x1 = grid->inner_end_x - grid->start_x;
    const int y0 = grid->inner_start_y - grid->start_y, y1 = grid->inner_end_y - grid->start_y;
    for (int y = y0; y < y1; ++y) {
        for (int x = x0; x < x1; ++x)
            out << "(" << p_real[(size_t)y * grid->dim_x + x] << "," << p_imag[(size_t)y * grid->dim_x + x] << ") ";
        out << std::endl;
    }
}

void State::write_particle_density(string prefix) {
    double *d = get_particle_density();
    const int w = grid->inner_end_x - grid->inner_start_x, h = grid->inner_end_y - grid->inner_start_y;
    write_matrix(prefix + "-density", d, w, w, h);
    delete[] d;
}

void State::write_phase(string prefix) {
    double *p = get_phase();
    const int w = grid->inner_end_x - grid->inner_start_x, h = grid->inner_end_y - grid->inner_start_y;
    write_matrix(prefix + "-phase", p, w, w, h);
    delete[] p;
}

// ---------------------------------------------------------------------------
// built-in states (src/model.cpp:640-748); on a Lattice1D the functions are
// evaluated at y = 0 with n_y = 0 (and omega_y = 1 for the Gaussian)
// ---------------------------------------------------------------------------
namespace {
template <typename F>
void fill_row(Lattice *grid, double *re, double *im, F f) {
    double xr, yr;
    for (int x = 0; x < grid->dim_x; ++x) {
        map_lattice_to_coordinate_space(grid, x, 0, &xr, &yr);
        const complex<double> v = f(xr, 0.);
        re[x] = real(v);
        im[x] = imag(v);
    }
}
}  // namespace

ExponentialState::ExponentialState(Lattice1D *g, int nx, double nrm, double ph, double *re, double *im)
    : State(g, 0, re, im), n_x(nx), n_y(0), norm(nrm), phase(ph) {
    fill_row(grid, p_real, p_imag, [this](double x, double y) { return exp_state(x, y); });
}

GaussianState::GaussianState(Lattice1D *g, double ox, double mx, double nrm, double ph, double *re, double *im)
    : State(g, 0, re, im), mean_x(mx), mean_y(0.), omega_x(ox), omega_y(1.), norm(nrm), phase(ph) {
    fill_row(grid, p_real, p_imag, [this](double x, double y) { return gauss_state(x, y); });
}

SinusoidState::SinusoidState(Lattice1D *g, int nx, double nrm, double ph, double *re, double *im)
    : State(g, 0, re, im), n_x(nx), n_y(0), norm(nrm), phase(ph) {
    fill_row(grid, p_real, p_imag, [this](double x, double y) { return sinusoid_state(x, y); });
}

ExponentialState::ExponentialState(Lattice2D *g, int nx, int ny, double nrm, double ph, double *re, double *im)
    : State(g, 0, re, im), n_x(nx), n_y(ny), norm(nrm), phase(ph) {
    double xr, yr;
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
            const complex<double> v = exp_state(xr, yr);
            p_real[(size_t)y * grid->dim_x + x] = real(v);
            p_imag[(size_t)y * grid->dim_x + x] = imag(v);
        }
}

complex<double> ExponentialState::exp_state(double x, double y) {
    const double lx = grid->global_no_halo_dim_x * grid->delta_x, ly = grid->global_no_halo_dim_y * grid->delta_y;
    return std::sqrt(norm / (lx * ly)) * std::exp(complex<double>(0., phase)) *
           std::exp(complex<double>(0., 2 * M_PI * double(n_x) / lx * x + 2 * M_PI * double(n_y) / ly * y));
}

GaussianState::GaussianState(Lattice2D *g, double ox, double oy, double mx, double my, double nrm, double ph,
                             double *re, double *im)
    : State(g, 0, re, im), mean_x(mx), mean_y(my), omega_x(ox), omega_y(oy), norm(nrm), phase(ph) {
    if (omega_y == -1.) omega_y = omega_x;
    double xr, yr;
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
            const complex<double> v = gauss_state(xr, yr);
            p_real[(size_t)y * grid->dim_x + x] = real(v);
            p_imag[(size_t)y * grid->dim_x + x] = imag(v);
        }
}

complex<double> GaussianState::gauss_state(double x, double y) {
    const double a = std::sqrt(norm * std::sqrt(omega_x * omega_y) / M_PI) *
                     std::exp(-(omega_x * std::pow(x - mean_x, 2.0) + omega_y * std::pow(y - mean_y, 2.0)) * 0.5);
    return complex<double>(a, 0.) * std::exp(complex<double>(0., phase));
}

SinusoidState::SinusoidState(Lattice2D *g, int nx, int ny, double nrm, double ph, double *re, double *im)
    : State(g, 0, re, im), n_x(nx), n_y(ny), norm(nrm), phase(ph) {
    double xr, yr;
    for (int y = 0; y < grid->dim_y; ++y)
        for (int x = 0; x < grid->dim_x; ++x) {
            map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
            const complex<double> v = sinusoid_state(xr, yr);
            p_real[(size_t)y * grid->dim_x + x] = real(v);
            p_imag[(size_t)y * grid->dim_x + x] = imag(v);
        }
}

complex<double> SinusoidState::sinusoid_state(double x, double y) {
    const double lx = grid->global_no_halo_dim_x * grid->delta_x, ly = grid->global_no_halo_dim_y * grid->delta_y;
    return std::sqrt(norm / (lx * ly)) * 2. * std::exp(complex<double>(0., phase)) *
           complex<double>(std::sin(2 * M_PI * double(n_x) / lx * x) * std::sin(2 * M_PI * double(n_y) / ly * y), 0.0);
}

// ---------------------------------------------------------------------------
// potentials (src/model.cpp:826-926)
// ---------------------------------------------------------------------------
Potential::Potential(Lattice *g, char *filename) : grid(g) {
    matrix = new double[(size_t)grid->dim_x * grid->dim_y];
    self_init = true;
    is_static = true;
    updated_potential_matrix = false;
    current_evolution_time = 0;
    static_potential = 0;
    evolving_potential = 0;
    std::ifstream in(filename);
    if (!in) throw std::runtime_error(std::string("cannot open ") + filename);
    for (size_t k = 0; k < (size_t)grid->dim_x * grid->dim_y; ++k) in >> matrix[k];
}

Potential::Potential(Lattice *g, double *ext) : grid(g) {
    self_init = ext == 0;
    matrix = ext ? ext : new double[(size_t)grid->dim_x * grid->dim_y]();
    is_static = true;
    updated_potential_matrix = false;
    current_evolution_time = 0;
    static_potential = 0;
    evolving_potential = 0;
}

Potential::Potential(Lattice *g, double (*f)(double, double)) : grid(g) {
    matrix = 0;
    self_init = false;
    is_static = true;
    updated_potential_matrix = false;
    current_evolution_time = 0;
    static_potential = f;
    evolving_potential = 0;
}

Potential::Potential(Lattice *g, double (*f)(double, double, double), int t) : grid(g) {
    matrix = 0;
    self_init = false;
    is_static = false;
    updated_potential_matrix = false;
    current_evolution_time = t;
    static_potential = 0;
    evolving_potential = f;
}

Potential::~Potential() {
    if (self_init) delete[] matrix;
}

double Potential::get_value(int x, int y) {
    if (matrix) return matrix[(size_t)y * grid->dim_x + x];
    double xr, yr;
    map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
    return is_static ? static_potential(xr, yr) : evolving_potential(xr, yr, current_evolution_time);
}

// src/model.cpp:894-902
bool Potential::update(double t) {
    if (current_evolution_time != t) {
        current_evolution_time = t;
        if (!is_static || updated_potential_matrix) return true;
    }
    return false;
}

HarmonicPotential::HarmonicPotential(Lattice2D *g, double ox, double oy, double m, double mx, double my)
    : Potential(g, const_potential), omegax(ox), omegay(oy), mass(m), mean_x(mx), mean_y(my) {}

HarmonicPotential::~HarmonicPotential() {}

double HarmonicPotential::get_value(int x, int y) {
    double xr, yr;
    map_lattice_to_coordinate_space(grid, x, y, &xr, &yr);
    xr -= mean_x;
    yr -= mean_y;
    return 0.5 * mass * (omegax * omegax * xr * xr + omegay * omegay * yr * yr);
}

// ---------------------------------------------------------------------------
// Hamiltonian (src/model.cpp:931-977)
// ---------------------------------------------------------------------------
Hamiltonian::Hamiltonian(Lattice *g, Potential *pot, double m, double ca, double clhy, double omega, double rx,
                         double ry)
    : mass(m), coupling_a(ca), LeeHuangYang_coupling_a(clhy), angular_velocity(omega), grid(g) {
    if (angular_velocity != 0.)
        throw std::runtime_error("A rotating frame (angular velocity != 0) is outside the HIP path.");
    rot_coord_x = (grid->global_dim_x - grid->periods[1] * 2 * grid->halo_x) * 0.5 + rx / grid->delta_x;
    rot_coord_y = (grid->global_dim_y - grid->periods[0] * 2 * grid->halo_y) * 0.5 + ry / grid->delta_y;
    self_init = pot == 0;
    potential = pot ? pot : new Potential(grid, const_potential);
}

Hamiltonian::~Hamiltonian() {
    if (self_init) delete potential;
}

// src/model.cpp:979-997: no LHY term; the second potential defaults to the first
Hamiltonian2Component::Hamiltonian2Component(Lattice *g, Potential *pot, Potential *pot_b, double m, double m_b,
                                             double ca, double cab, double cb, double om_r, double om_i,
                                             double omega, double rx, double ry)
    : Hamiltonian(g, pot, m, ca, 0., omega, rx, ry), mass_b(m_b), coupling_ab(cab), coupling_b(cb), omega_r(om_r),
      omega_i(om_i) {
    potential_b = pot_b ? pot_b : potential;
}

Hamiltonian2Component::~Hamiltonian2Component() {}
