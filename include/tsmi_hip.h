/*
 * tsmi_hip.h -- C ABI of the MI355X (gfx950) Trotter-Suzuki step library
 * (libtsmi_hip.so).
 *
 * This is the drop-in boundary for the reference's kernel interface
 * ITrotterKernel (src/trottersuzuki.h:502-520): a host-side kernel object
 * (src/kernel.h:66-151 for the CPU one) becomes an opaque context that owns the
 * device copies of the wave function, of the potential factor E = exp(-i dt V)
 * and of the raw potential.  Only plain C types cross the boundary: pointers,
 * sizes, doubles and int32.  Every call returns TSMI_OK (0) or an error code;
 * tsmi_last_error() gives the message of the calling thread's last failure
 * (the reference reports errors through my_abort -> std::runtime_error,
 * src/common.cpp:104-115; the C++ wrapper turns a non-zero status into that).
 *
 * Arrays passed by the host use the reference tile layout: row-major W x H
 * doubles, idx = y*W + x, with W = dim_x + 8*periodic_x and H = dim_y +
 * 8*periodic_y (src/model.cpp:140-157, src/common.cpp:83-102).
 */
#ifndef TSMI_HIP_H
#define TSMI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TSMI_OK = 0,
    TSMI_E_INVALID = 1,      /* bad argument or state                         */
    TSMI_E_HIP = 2,          /* HIP runtime error (message has the HIP string) */
    TSMI_E_OOM = 3,          /* device allocation failed                      */
    TSMI_E_UNSUPPORTED = 4   /* feature outside this library's path           */
};

typedef struct tsmi_ctx tsmi_ctx;

/* Everything CPUBlock's constructor reads from Lattice, Hamiltonian and the
 * Solver (src/cpukernel.cpp:172-272, src/kernel.h:68-70), for one component. */
typedef struct {
    int32_t dim_x, dim_y;          /* inner lattice points (global_no_halo_dim_x/y)      */
    int32_t periodic_x, periodic_y;
    int32_t halo_x, halo_y;        /* lattice halo widths: 4 (halo_y 0 for a one-row chain) */
    int32_t imag_time;             /* 0: real time, 1: imaginary time                    */
    int32_t precision;             /* 64: FP64 throughout; 32: FP32 bulk, FP64 reductions */
    int32_t tile_rows;             /* rows per wavefront segment; 0 chooses              */
    int32_t sweeps;                /* steps per pass over HBM in tsmi_evolve: 0 chooses
                                      (2 where exact), 1, or 2 (real time with both axes
                                      closed or both periodic only)                      */
    double delta_t, mass;
    double coupling;               /* g (Hamiltonian::coupling_a)                        */
    double coupling_lhy;           /* g_LHY (Hamiltonian::LeeHuangYang_coupling_a)       */
    double delta_x, delta_y;
    double length_x, length_y;
    double norm_target;            /* imaginary time: squared norm kept (0: none)        */
} tsmi_desc;

/* Creates a context on the current HIP device: the CPUBlock/CC2Kernel
 * constructor (src/kernel.h:68-70, src/gpukernel.cu:122-238).  Rejects angular
 * velocity/two components/cylindrical by construction (not in the descriptor). */
int tsmi_create(const tsmi_desc *desc, tsmi_ctx **out);

/* Releases the context (~CPUBlock, src/cpukernel.cpp:399-411). */
void tsmi_destroy(tsmi_ctx *ctx);

/* Runs the context's work on an existing HIP stream (hipStream_t passed as a
 * pointer; NULL is the HIP null stream).  Until this is called the context uses
 * a non-blocking stream of its own.  No reference counterpart: the reference's
 * GPU kernel creates its own streams. */
int tsmi_set_stream(tsmi_ctx *ctx, void *hip_stream);

/* Uploads the wave function, W*H doubles each: the state buffers the kernel is
 * constructed on (CPUBlock borrows state->p_real/p_imag, src/cpukernel.cpp:244-245). */
int tsmi_set_state(tsmi_ctx *ctx, const double *re, const double *im);

/* The same, from device memory: re/im point to W*H FP64 arrays (row stride W)
 * on the context's device.  Asynchronous on the context stream. */
int tsmi_set_state_device(tsmi_ctx *ctx, const double *re, const double *im);

/* Uploads E = exp(-i dt V) (real time) or exp(-dt V) (imaginary time, e_im may
 * be NULL), W*H doubles, rounded once to the context precision:
 * ITrotterKernel::update_potential (src/trottersuzuki.h:514) and the E tables
 * given to the constructor (src/solver.cpp:74-109,124). */
int tsmi_set_exp_potential(tsmi_ctx *ctx, const double *e_re, const double *e_im);

/* The same, from device memory (W*H FP64, row stride W).  Asynchronous. */
int tsmi_set_exp_potential_device(tsmi_ctx *ctx, const double *e_re, const double *e_im);

/* Uploads the raw potential V, W*H doubles, used by tsmi_energy only
 * (Potential::get_value in Solver::calculate_energy_expected_values,
 * src/solver.cpp:336). */
int tsmi_set_potential(tsmi_ctx *ctx, const double *v);

/* The same, from device memory (W*H FP64, row stride W).  Asynchronous. */
int tsmi_set_potential_device(tsmi_ctx *ctx, const double *v);

/* Tabulates E from a raw potential V already on the device (W*H FP64, row
 * stride W): E = exp(-i dt V) in real time, exp(-dt V) in imaginary time, as
 * Solver::initialize_exp_potential computes it on the host (src/solver.cpp:74-109),
 * and keeps V for tsmi_energy.  This is the time-dependent-potential update of
 * Solver::evolve (src/solver.cpp:187-192: Potential::update, re-tabulation,
 * ITrotterKernel::update_potential) with the table built on the device.
 * Asynchronous on the context stream. */
int tsmi_set_potential_exp_device(tsmi_ctx *ctx, const double *v);

/* n_steps full steps, as the single-component loop of Solver::evolve
 * (src/solver.cpp:186-230): per step run_kernel_on_halo + run_kernel
 * (src/cpukernel.cpp:413-524), the periodic halo exchange except after the last
 * step (:691-747), and in imaginary time the renormalisation of
 * wait_for_completion (:551-570).  Asynchronous on the context stream. */
int tsmi_evolve(tsmi_ctx *ctx, int32_t n_steps);

/* n_steps consecutive run_kernel calls as Solver::evolve issues them
 * (src/solver.cpp:201-210): each step followed by the periodic halo exchange,
 * the last one only when exchange_last != 0.  tsmi_evolve(n) is
 * tsmi_run_steps(n, 0); a kernel object that queues the Solver's per-step calls
 * issues them with this, so they reach the two-step pass.  Asynchronous. */
int tsmi_run_steps(tsmi_ctx *ctx, int32_t n_steps, int32_t exchange_last);

/* One step, followed by the periodic halo exchange when exchange != 0:
 * run_kernel_on_halo + run_kernel with start/finish_halo_exchange
 * (src/cpukernel.cpp:413-524, :691-747) and, in imaginary time,
 * wait_for_completion's renormalisation (:551-570).  Asynchronous. */
int tsmi_step(tsmi_ctx *ctx, int32_t exchange);

/* Gradient of a loss of the final state with respect to the potential.
 * No reference counterpart (the reference evolves forward only).  Real time
 * only (TSMI_E_UNSUPPORTED otherwise); any g and g_LHY; closed or periodic axes,
 * except a periodic x axis with a closed y axis of a 2D lattice (the reference
 * refreshes the x halo of only H - 2*halo_y rows there) and a periodic axis
 * whose inner extent is below twice the halo.
 *
 * tsmi_evolve_checkpointed: tsmi_evolve(n_steps) that also keeps the state
 * every `every` steps on the device (ceil(n_steps/every) copies of the tile).
 * Asynchronous.  Any later evolve, step or set_* call drops the checkpoints.
 *
 * tsmi_adjoint / tsmi_adjoint_device: the reverse sweep over those steps.
 * g_re/g_im = dL/dRe psi_N and dL/dIm psi_N (W*H FP64); out grad_v = dL/dV on
 * every tile cell for E = exp(-i dt V) (W*H FP64); g0_re/g0_im (optional, may
 * be NULL) = dL/dRe psi_0 and dL/dIm psi_0.  Each reverse step is the step at
 * -dt applied to the adjoint and to the state together, the state re-anchored
 * on the checkpoints; the FP64 gradient accumulates dt Im(conj(mu) y) of the
 * two mid-step states.  On periodic axes the lattice is the periodic one: the
 * halo entries of g are taken as acting on the inner cells they mirror, and
 * grad_v and g0 hold, on each inner cell, the derivative with respect to the
 * value the cell shares with its halo copies (zero on the halo).  The host form
 * syncs; the device form is asynchronous on the context stream.  The state of
 * the context is left at psi_N. */
int tsmi_evolve_checkpointed(tsmi_ctx *ctx, int32_t n_steps, int32_t every);
int tsmi_adjoint(tsmi_ctx *ctx, const double *g_re, const double *g_im, double *grad_v,
                 double *g0_re, double *g0_im);
int tsmi_adjoint_device(tsmi_ctx *ctx, const double *g_re, const double *g_im, double *grad_v,
                        double *g0_re, double *g0_im);

/* Batches: many independent lattices advanced together.  No reference
 * counterpart (the reference runs one Solver per lattice).  Each lattice has its
 * own descriptor (size, boundaries, dt, mass, spacing, couplings); precision,
 * time direction and whether g_LHY != 0 are common to the batch.  Each lattice
 * gets exactly tsmi_evolve's steps.  A lattice whose descriptor allows two steps
 * per pass (sweeps 0 or 2, real time, both axes closed or both periodic) is
 * advanced like a context with that setting; the others one step per pass.  A
 * pass of the batch is one persistent launch over all lattices of its kind
 * (plus one halo-exchange launch when some of them are periodic).
 *
 * tsmi_batch_evolve(n) records its launch sequence in a HIP graph the first time
 * n is used from given buffer parities and replays the graph afterwards;
 * asynchronous.  tsmi_batch_tune times the batch (`steps` steps per trial, at
 * least 6 when some lattice can take two steps per pass, at least 24 when some
 * lattice can take either) for candidate segment
 * heights of each size class's dominant pass and, for lattices whose descriptor
 * leaves the choice open (sweeps 0), for one and for two steps per pass; one
 * class (kind and tile shape) at a time; keeps the fastest, restores the states
 * and reports the tuned time per step.  tsmi_batch_info: out[0..2] W, H, pitch
 * of lattice i's tile, out[3..5] strips, segments and rows per segment of its
 * dominant pass, out[6] steps per pass (1 or 2).
 * set_state / set_exp_potential / get_state take W*H FP64 host arrays of
 * lattice i; get_state syncs. */
typedef struct tsmi_batch tsmi_batch;
int tsmi_batch_create(const tsmi_desc *descs, int32_t n, tsmi_batch **out);
void tsmi_batch_destroy(tsmi_batch *batch);
int tsmi_batch_set_state(tsmi_batch *batch, int32_t i, const double *re, const double *im);
int tsmi_batch_set_exp_potential(tsmi_batch *batch, int32_t i, const double *e_re, const double *e_im);
int tsmi_batch_evolve(tsmi_batch *batch, int32_t n_steps);
int tsmi_batch_get_state(tsmi_batch *batch, int32_t i, double *re, double *im);
int tsmi_batch_tune(tsmi_batch *batch, int32_t steps, double *ms_per_step);
int tsmi_batch_info(const tsmi_batch *batch, int32_t i, int32_t out[7]);
int tsmi_batch_synchronize(tsmi_batch *batch);

/* Two-component systems: Hamiltonian2Component (src/trottersuzuki.h:459-497)
 * evolved by the two-state CPUBlock (src/cpukernel.cpp:274-392) under the
 * two-state Solver (src/solver.cpp:43-60,150-242).
 *
 * a and b describe the two components on one lattice: every lattice, time-step,
 * time-direction, precision and tile_rows field identical; a.mass / b.mass =
 * Hamiltonian2Component::mass / mass_b; a.coupling / b.coupling = coupling_a /
 * coupling_b; coupling_lhy 0 (the two-state kernel has no LHY term,
 * src/cpukernel.cpp:332-335); a.norm_target / b.norm_target = the squared norms
 * of the two states when the kernel is built (src/solver.cpp:153-159), the
 * imaginary-time targets.  a.sweeps picks the steps per pass of tsmi_pair_evolve:
 * 0 chooses two wherever they are exact (FP32, real time, both axes closed or both
 * periodic, and both E tables of unit modulus; otherwise one), 1 forces one, 2
 * requires a lattice and precision where two can run.
 * b.sweeps is not read. */
typedef struct {
    tsmi_desc a, b;
    double coupling_ab;              /* Hamiltonian2Component::coupling_ab      */
    double omega_r, omega_i;         /* Hamiltonian2Component::omega_r, omega_i */
} tsmi_pair_desc;

typedef struct tsmi_pair tsmi_pair;

/* The two-state CPUBlock constructor (src/cpukernel.cpp:274-392). */
int tsmi_pair_create(const tsmi_pair_desc *desc, tsmi_pair **out);
void tsmi_pair_destroy(tsmi_pair *pair);

/* Per component (which = 0: a, 1: b), W*H FP64 host arrays: the state
 * (CPUBlock borrows state1/state2, src/cpukernel.cpp:364-367), E = exp(-i dt V)
 * or exp(-dt V) (external_pot_real/imag[which], src/solver.cpp:74-109,
 * ITrotterKernel::update_potential(..., which)), the raw potential for
 * tsmi_pair_energy. */
int tsmi_pair_set_state(tsmi_pair *pair, int32_t which, const double *re, const double *im);
int tsmi_pair_set_exp_potential(tsmi_pair *pair, int32_t which, const double *e_re, const double *e_im);
int tsmi_pair_set_potential(tsmi_pair *pair, int32_t which, const double *v);

/* The two-component loop of Solver::evolve (src/solver.cpp:172-230): the half
 * Rabi rotation first when initial_half != 0 (the C++ API always; the Python
 * API only at t = 0, :174-175), then per step both components' steps (each
 * potential pass sees the other component's start-of-step density), the halo
 * exchange of both except after the last step, in imaginary time each
 * component's renormalisation, the Rabi rotation by 1 (by 1/2 after the last
 * step) and in imaginary time with a Rabi coupling the joint renormalisation
 * (src/cpukernel.cpp:551-667).  The norm targets persist between calls, as in a
 * kernel the Solver keeps.  Asynchronous. */
int tsmi_pair_evolve(tsmi_pair *pair, int32_t n_steps, int32_t initial_half);

/* The same loop one piece at a time, for a kernel object driven by the
 * reference's Solver through ITrotterKernel (src/trottersuzuki.h:502-520).
 * tsmi_pair_step: both components' steps (run_kernel_on_halo + run_kernel +
 * wait_for_completion of each, src/cpukernel.cpp:413-570), the halo exchange of
 * both when exchange != 0, and in real time the following rabi_coupling(rabi_var)
 * applied as the rows leave the step (rabi_var = 0: none).  tsmi_pair_rabi:
 * rabi_coupling(var) of both current tiles (src/cpukernel.cpp:581-608) and, when
 * normalize != 0 in imaginary time, normalization() (:610-667).  Asynchronous. */
int tsmi_pair_step(tsmi_pair *pair, int32_t exchange, double rabi_var);
int tsmi_pair_rabi(tsmi_pair *pair, double var, int32_t normalize);

/* get_sample of one component, the whole tile (src/cpukernel.cpp:572-579); syncs. */
int tsmi_pair_get_state(tsmi_pair *pair, int32_t which, double *re, double *im);

/* Solver::calculate_energy_expected_values for two components
 * (src/solver.cpp:244-564, angular velocity 0): out[0..11] = total, kinetic a,
 * kinetic b, potential a, potential b, intra a, intra b, inter-species, Rabi,
 * squared norm a, squared norm b, total squared norm.  Needs both raw
 * potentials; syncs. */
int tsmi_pair_energy(tsmi_pair *pair, double out[12]);

/* out[0..5] = W, H, pitch, strips, segments, rows per segment of the pair step. */
int tsmi_pair_geometry(const tsmi_pair *pair, int32_t out[6]);
int tsmi_pair_synchronize(tsmi_pair *pair);

/* Kernel timing for measurement (no reference counterpart).  With timing on,
 * every step-kernel launch is bracketed by HIP events on the context stream;
 * tsmi_kernel_time syncs and reports out[0..3] = total ms and number of the
 * fused two-step launches, total ms and number of the single-step launches
 * since timing was (re)enabled. */
int tsmi_kernel_timing(tsmi_ctx *ctx, int32_t enable);
int tsmi_kernel_time(tsmi_ctx *ctx, double out[4]);

/* Squared norm of the inner region times dx*dy:
 * ITrotterKernel::calculate_squared_norm (src/cpukernel.cpp:526-549).  Syncs. */
int tsmi_squared_norm(tsmi_ctx *ctx, double *out);

/* Energies of the current state, as Solver::calculate_energy_expected_values
 * computes them for one component (src/solver.cpp:244-564):
 * out[0..5] = total, kinetic, potential, intra-species, LHY, squared norm.
 * Needs tsmi_set_potential.  Syncs. */
int tsmi_energy(tsmi_ctx *ctx, double out[6]);

/* The energies of tsmi_energy for a state held in the caller's device memory,
 * allocating nothing.  No counterpart in the reference: for callers that
 * cannot allocate device memory -- code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc.  desc describes the lattice as for
 * tsmi_create (precision 32: the planes hold floats, 64: doubles); psi_re,
 * psi_im: the W x H tile (halo included on periodic axes) with a row pitch of
 * `pitch` elements (pitch >= W); v: the raw potential as W x H doubles with the
 * same pitch, or NULL (potential term 0); out: 6 doubles in device memory,
 * written as tsmi_energy's out[0..5].  The block partials go to a fixed device
 * array of the library (TSMI_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held
 * until the process ends, every element written before it is read); calls that
 * overlap on one GPU race on it.  The work is queued on `stream` (a
 * hipStream_t; NULL: the null stream) and not waited for.  Inner regions wider
 * than TSMI_NOALLOC_MAX_COLUMNS columns are rejected (TSMI_E_INVALID). */
#define TSMI_NOALLOC_SCRATCH_BYTES (4096 * 14 * 8)
#define TSMI_NOALLOC_MAX_COLUMNS (4096 * 240)
int tsmi_energy_device_noalloc(const tsmi_desc *desc, const void *psi_re, const void *psi_im, int64_t pitch,
                               const double *v, double *out, void *stream);

/* State::calculate_expected_values (src/model.cpp:385-536):
 * out[0..9] = norm2, <X>, <X^2>, <Y>, <Y^2>, <P_x>, <P_x^2>, <P_y>, <P_y^2>, <L_z>.
 * Syncs. */
int tsmi_expected_values(tsmi_ctx *ctx, double out[10]);

/* Copies a width x height window at tile position (x, y) into host arrays with
 * row stride dest_stride: ITrotterKernel::get_sample (src/trottersuzuki.h:508,
 * src/cpukernel.cpp:572-579), one component.  Syncs. */
int tsmi_get_sample(tsmi_ctx *ctx, size_t dest_stride, size_t x, size_t y,
                    size_t width, size_t height, double *dest_re, double *dest_im);

/* The whole tile into device memory: W*H FP64 arrays (row stride W), with a
 * pending imaginary-time renormalisation applied (get_sample into device
 * buffers).  Asynchronous on the context stream. */
int tsmi_get_state_device(tsmi_ctx *ctx, double *re, double *im);

/* Waits for the context's queued work (the CUDA kernel's wait_for_completion
 * synchronisation, src/gpukernel.cu:502-521). */
int tsmi_synchronize(tsmi_ctx *ctx);

/* Tile geometry chosen for the context: out[0..5] = W, H, pitch (elements),
 * strips, segments, rows per segment.  No reference counterpart (reporting). */
int tsmi_geometry(const tsmi_ctx *ctx, int32_t out[6]);

/* The geometry tsmi_create would choose for a descriptor, without creating a
 * context or touching device memory (with no device to query, the segment
 * heights assume an unknown occupancy).  out[0..11] = W, H, pitch (elements),
 * strips, segments and rows per segment of the one-step pass, steps per pass
 * of tsmi_evolve (1 or 2), strips, segments and rows per segment of the
 * two-step pass (0 when it is not used), and the allocation's width and height
 * (the 8-wide halo layout of a doubly periodic lattice at two steps per pass).
 * Every segment plus its apron rows spans less than 2^31 bytes of an FP64
 * plane.  No reference counterpart (planning and reporting). */
int tsmi_plan_geometry(const tsmi_desc *desc, int32_t out[12]);

/* Message of the calling thread's last failed call ("" if none). */
const char *tsmi_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
