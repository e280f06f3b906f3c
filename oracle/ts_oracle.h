/*
 * ts_oracle.h -- CPU restatement of the reference Trotter-Suzuki step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and only there, as the checker.  The product library never calls it.
 *
 * Layout (same as the reference tile, src/trottersuzuki.h:40-51):
 *   SoA double re[W*H], im[W*H], row-major, idx = y*W + x, where W and H include
 *   the 4-cell halo on periodic axes (src/model.cpp:140-157, src/common.cpp:83-102).
 *   Coordinates (x, y) are tile-local; the checkerboard parity uses them.
 *
 * Pinning status of each function is stated beside it (see the tests in
 * tests/test_oracle_*.py).
 */
#ifndef TS_ORACLE_H
#define TS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int W, H;               /* tile width/height incl. halo                       */
    int halo_x, halo_y;     /* lattice halo widths (4 for the Cartesian path)     */
    int periodic_x, periodic_y;
    int imag_time;          /* 0: real time, 1: imaginary time                    */
    double aH, bH, aV, bV;  /* kinetic pair coefficients (cpukernel.cpp:195-208)  */
    double c_nl, c_lhy;     /* g*dt and g_LHY*dt (cpukernel.cpp:213,216)          */
    double dx, dy;          /* lattice spacings                                   */
    double norm_target;     /* imag time: squared norm kept by renormalisation    */
} ts_oracle_desc;

/* Coefficients exactly as CPUBlock's constructor computes them
 * (src/cpukernel.cpp:195-216).  out[0..5] = aH, bH, aV, bV, c_nl, c_lhy. */
void ts_oracle_coefficients(double dt, double mass, double dx, double dy,
                            double g, double g_lhy, int imag_time, double out[6]);

/* E = exp(-i dt V) (real time) or exp(-dt V) (imaginary time) on n cells, with
 * the complex exponential of the C library as Solver::initialize_exp_potential
 * uses it (src/solver.cpp:98-105).  Pinned: closed form |E|=1 / E=exp(-dt V). */
void ts_oracle_exp_potential(long long n, const double *v, double dt, int imag_time,
                             double *e_re, double *e_im);

/* One full second-order step V0 H0 V1 H1 P H1 V1 H0 V0 on the whole tile with
 * closed tile edges, in place (src/cpukernel.cpp:22-92, src/cpucartesian.cpp:4-138).
 * e_im is ignored in imaginary time.  Pinned: golden vectors + dense brute force. */
void ts_oracle_full_step(const ts_oracle_desc *d, double *re, double *im,
                         const double *e_re, const double *e_im);

/* The individual sub-passes, exposed so the tests can check each against its
 * closed form.  s is the pass parity (0 or 1). */
void ts_oracle_pass_vertical(const ts_oracle_desc *d, int s, double a, double b, double *re, double *im);
void ts_oracle_pass_horizontal(const ts_oracle_desc *d, int s, double a, double b, double *re, double *im);
void ts_oracle_pass_potential(const ts_oracle_desc *d, double *re, double *im,
                              const double *e_re, const double *e_im);

/* Periodic halo refresh of the single-process build
 * (src/cpukernel.cpp:707-715 then :738-746).  Pinned: golden vectors. */
void ts_oracle_halo_exchange(const ts_oracle_desc *d, double *re, double *im);

/* Squared norm over the inner region times dx*dy (src/cpukernel.cpp:526-549). */
double ts_oracle_squared_norm(const ts_oracle_desc *d, const double *re, const double *im);

/* State::get_squared_norm: the same inner sum scaled as (dx*dy)
 * (src/model.cpp:418,534); it is the imaginary-time target of Solver::evolve
 * (src/solver.cpp:155). */
double ts_oracle_state_norm(const ts_oracle_desc *d, const double *re, const double *im);

/* Solver::evolve(n) for one component on a freshly built kernel
 * (src/solver.cpp:186-239): per step the full step, the halo exchange except
 * after the last step, and in imaginary time the renormalisation of the whole
 * tile (src/cpukernel.cpp:551-563).  In place.  Pinned: golden vectors. */
void ts_oracle_evolve(const ts_oracle_desc *d, int n_steps, double *re, double *im,
                      const double *e_re, const double *e_im);

/* Energies of one component (src/solver.cpp:244-564 with a single component and
 * angular velocity 0).  v is the raw potential on the tile.  Lx is the physical
 * side along x (used by the returned norm, src/solver.cpp:561).
 * out[0..5] = total, kinetic, potential, intra-species, LHY, squared norm. */
void ts_oracle_energy(const ts_oracle_desc *d, const double *re, const double *im,
                      const double *v, double mass, double g, double g_lhy,
                      double Lx, int Nx, double out[6]);

/* State::calculate_expected_values (src/model.cpp:385-536).  x_coord[W] and
 * y_coord[H] are the mapped physical coordinates of the tile columns/rows.
 * out[0..9] = norm2, <X>, <X^2>, <Y>, <Y^2>, <Px>, <Px^2>, <Py>, <Py^2>, <Lz>. */
void ts_oracle_expected_values(const ts_oracle_desc *d, const double *re, const double *im,
                               const double *x_coord, const double *y_coord, double out[10]);

/* Gradient of a loss L(psi_N) with respect to the potential V on every tile
 * cell, through n_steps real-time steps of ts_oracle_evolve (closed axes,
 * g = g_LHY = 0): the discrete adjoint, not a reference function.
 * g_re/g_im = dL/dRe psi_N, dL/dIm psi_N; E = exp(-i dt V) as for the forward.
 * Out: grad_v[W*H] = dL/dV; g0_re/g0_im (may be NULL) = dL/dRe psi_0, dL/dIm psi_0.
 * Checkpoints every `every` steps, each segment recomputed forward.
 * Returns 0, -1 for an unsupported descriptor, -2 when out of memory.
 * Pinned: central finite differences of ts_oracle_evolve (tests/test_adjoint.py). */
int ts_oracle_adjoint(const ts_oracle_desc *d, int n_steps, int every,
                      const double *psi_re, const double *psi_im,
                      const double *e_re, const double *e_im, double dt,
                      const double *g_re, const double *g_im,
                      double *grad_v, double *g0_re, double *g0_im);

#ifdef __cplusplus
}
#endif
#endif
