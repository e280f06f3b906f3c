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
 * cell, through n_steps real-time steps of ts_oracle_evolve (any boundaries,
 * g and g_LHY included): the discrete adjoint, not a reference function.  Each
 * tile entry of V and of psi_0 is its own variable (a halo entry included).
 * g_re/g_im = dL/dRe psi_N, dL/dIm psi_N; E = exp(-i dt V) as for the forward.
 * Out: grad_v[W*H] = dL/dV; g0_re/g0_im (may be NULL) = dL/dRe psi_0, dL/dIm psi_0.
 * Checkpoints every `every` steps, each segment recomputed forward.
 * Returns 0, -1 for an unsupported descriptor (imaginary time), -2 when out of
 * memory.  Pinned: central finite differences of ts_oracle_evolve
 * (tests/test_adjoint_oracle.py). */
int ts_oracle_adjoint(const ts_oracle_desc *d, int n_steps, int every,
                      const double *psi_re, const double *psi_im,
                      const double *e_re, const double *e_im, double dt,
                      const double *g_re, const double *g_im,
                      double *grad_v, double *g0_re, double *g0_im);

/* Adjoint (transpose) of ts_oracle_halo_exchange, in place: every halo entry is
 * added onto the cell it copies and cleared.  On a field of derivatives with
 * respect to tile entries it gives the derivatives with respect to the values
 * shared by a cell and its halo copies (a periodic field). */
void ts_oracle_halo_exchange_adjoint(const ts_oracle_desc *d, double *re, double *im);


/* ---------------------------------------------------------------------------
 * Two-component systems (Hamiltonian2Component, src/trottersuzuki.h:459-497,
 * and the two-state CPUBlock, src/cpukernel.cpp:274-392).
 * ------------------------------------------------------------------------- */
typedef struct {
    ts_oracle_desc a, b;     /* same lattice, time direction and spacings; each its own
                                aH..bV (mass, mass_b) and c_nl = dt*g_a, dt*g_b
                                (src/cpukernel.cpp:294-328); c_lhy must be 0 (the
                                two-state potential pass has no LHY term, :332-335,
                                src/cpucartesian.cpp:78-82)                         */
    double c_ab;             /* dt*g_ab (src/cpukernel.cpp:329)                      */
    double omega_r, omega_i; /* Hamiltonian2Component::omega_r/_i; the kernel uses
                                half of each (src/cpukernel.cpp:330-331)             */
} ts_oracle_desc2;

/* CPUBlock::rabi_coupling(var, dt) (src/cpukernel.cpp:581-608) on the whole
 * tiles, in place (src/cpucartesian.cpp:264-290).  Pinned: the closed-form
 * population transfer cos^2(omega t / 2) and golden vectors. */
void ts_oracle_rabi(const ts_oracle_desc2 *d, double var, double dt,
                    double *a_re, double *a_im, double *b_re, double *b_im);

/* Solver::evolve(n) of a two-component system on one kernel that persists
 * between calls (src/solver.cpp:150-242): the half Rabi rotation first when
 * initial_half != 0 (:173-177; the C++ API always, the Python one only at t = 0);
 * per step component a's full step, whose potential pass also sees the other
 * component's start-of-step density |psi_b|^2 (src/cpucartesian.cpp:75-92,
 * :115-126), its halo exchange except after the last step, its imaginary-time
 * renormalisation to norm[0] (src/cpukernel.cpp:551-570); component b likewise
 * with the start-of-step |psi_a|^2 and norm[1]; the Rabi rotation by var = 1
 * (0.5 after the last step, :222-225); in imaginary time with a Rabi coupling the joint
 * renormalisation to tot_norm, which also redistributes norm[]
 * (src/cpukernel.cpp:610-667).  norm[2] is the kernel's per-component target,
 * in/out; tot_norm = norm[0] + norm[1] at kernel construction
 * (src/cpukernel.cpp:321-325).
 * Pinned: golden vectors, closed-form Rabi transfer, the reference's
 * imaginary_mixed_BEC_test norms. */
void ts_oracle_evolve2(const ts_oracle_desc2 *d, int n_steps, int initial_half, double dt,
                       double *a_re, double *a_im, double *b_re, double *b_im,
                       const double *ea_re, const double *ea_im, const double *eb_re, const double *eb_im,
                       double norm[2], double tot_norm);

/* Energies of a two-component system (src/solver.cpp:244-564, angular velocity
 * 0).  va, vb are the raw potentials on the tile.  out[0..11] = total, kinetic
 * a, kinetic b, potential a, potential b, intra a, intra b, inter-species, Rabi,
 * squared norm a, squared norm b, total squared norm.  Pinned: golden vectors. */
void ts_oracle_energy2(const ts_oracle_desc2 *d, const double *a_re, const double *a_im,
                       const double *b_re, const double *b_im, const double *va, const double *vb,
                       double mass_a, double mass_b, double g_a, double g_b, double g_ab,
                       double omega_r, double omega_i, double Lx, int Nx, double out[12]);

#ifdef __cplusplus
}
#endif
#endif
