/*
 * hip_kernel.h -- HipKernel, the MI355X implementation of ITrotterKernel.
 *
 * HipKernel is a sibling of the reference's CPUBlock (src/kernel.h:66-151) with
 * the same constructor arguments (src/kernel.h:68-70), so the reference's kernel
 * factory (Solver::init_kernel, src/solver.cpp:118-148) can build it for a new
 * kernel type.  It compiles against the reference's own trottersuzuki.h or
 * against this project's include/trottersuzuki.h: it only uses the members both
 * declare.  All device work goes through the C ABI of libtsmi_hip.so
 * (include/tsmi_hip.h); failures throw std::runtime_error, as my_abort does in
 * a build without MPI (src/common.cpp:104-115).
 *
 * Supported: one component, Cartesian coordinates, angular velocity 0, real or
 * imaginary time, closed or periodic axes, FP64 or FP32 bulk.
 *
 * HipKernel queues the Solver's per-step calls (run_kernel, with or without the
 * halo exchange that start_halo_exchange asks for) and issues the queue as one
 * tsmi_run_steps call when something reads or changes the state: get_sample,
 * calculate_squared_norm, update_potential, the energies, or a step that follows
 * one without an exchange.  A run of steps therefore reaches the library's
 * two-step pass (two steps per HBM round trip where that is exact) with the
 * reference's halo contents at every call boundary.
 *
 * HipPairKernel is the sibling of the two-state CPUBlock constructor
 * (src/cpukernel.cpp:274-392) for a two-component system, with the same
 * arguments.  The Solver drives a step as run_kernel + wait_for_completion for
 * each component, then rabi_coupling(var) and normalization()
 * (src/solver.cpp:201-226); the device runs both components' steps as one pass,
 * so HipPairKernel queues the two run_kernel calls and issues the pass when
 * rabi_coupling supplies var (real time: the rotation is applied as the rows
 * leave the pass; imaginary time: the rotation and the joint renormalisation
 * follow as one more pass at normalization()).
 */
#ifndef TSMI_HIP_KERNEL_H
#define TSMI_HIP_KERNEL_H

#include <string>

// angle brackets: the include path decides which trottersuzuki.h is used
#include <trottersuzuki.h>
#include <tsmi_hip.h>

class HipKernel : public ITrotterKernel {
public:
    /** precision: 64 (FP64 throughout) or 32 (FP32 bulk, FP64 reductions). */
    HipKernel(Lattice *grid, State *state, Hamiltonian *hamiltonian, double *external_pot_real,
              double *external_pot_imag, double delta_t, double norm, bool imag_time, int precision = 64);
    ~HipKernel();

    void run_kernel();                 ///< queues one step; the halo exchange if start_halo_exchange() preceded it
    void run_kernel_on_halo();         ///< no-op: the step kernel covers the whole tile
    void wait_for_completion();        ///< no-op: the renormalisation is part of the step
    void get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height,
                    double *dest_real, double *dest_imag, double *dest_real2 = 0, double *dest_imag2 = 0) const;
    void normalization();              ///< two components only: throws
    void rabi_coupling(double var, double delta_t);   ///< two components only: throws
    double calculate_squared_norm(bool global = true) const;
    bool runs_in_place() const { return false; }
    std::string get_name() const { return precision_ == 32 ? "HIP-FP32" : "HIP"; }
    void update_potential(double *_external_pot_real, double *_external_pot_imag, int which);
    void cpy_first_positive_to_first_negative() {}    ///< cylindrical only: nothing to do
    void start_halo_exchange();        ///< marks the next run_kernel() as followed by the exchange
    void finish_halo_exchange() {}     ///< the exchange runs on the device right after the step

    // Beyond the interface: used by this project's Solver.
    void evolve(int n_steps);          ///< n steps, the exchange after all but the last
    void set_potential(const double *v);   ///< raw V on the tile, for energy()
    void energy(double out[6]) const;  ///< total, kinetic, potential, intra, LHY, squared norm
    void expected_values(double out[10]) const;
    void flush() const;                ///< issues the queued steps
    tsmi_ctx *context() const { return ctx_; }   ///< flush() before using it directly

private:
    tsmi_ctx *ctx_;
    int precision_;
    int width_, height_;
    bool exchange_next_;
    mutable int queued_;               ///< steps requested and not yet issued
    mutable bool queued_exchange_;     ///< whether the last of them is followed by the exchange
};

class HipPairKernel : public ITrotterKernel {
public:
    HipPairKernel(Lattice *grid, State *state1, State *state2, Hamiltonian2Component *hamiltonian,
                  double **external_pot_real, double **external_pot_imag, double delta_t, double *norm,
                  bool imag_time, int precision = 64);
    ~HipPairKernel();

    void run_kernel();                 ///< queues the current component's step
    void run_kernel_on_halo() {}       ///< the step covers the whole tile
    void wait_for_completion();        ///< moves to the other component
    void get_sample(size_t dest_stride, size_t x, size_t y, size_t width, size_t height,
                    double *dest_real, double *dest_imag, double *dest_real2 = 0, double *dest_imag2 = 0) const;
    void normalization();
    void rabi_coupling(double var, double delta_t);
    double calculate_squared_norm(bool global = true) const;   ///< of the current component
    bool runs_in_place() const { return false; }
    std::string get_name() const { return precision_ == 32 ? "HIP-FP32" : "HIP"; }
    void update_potential(double *_external_pot_real, double *_external_pot_imag, int which);
    void cpy_first_positive_to_first_negative() {}
    void start_halo_exchange();        ///< the queued steps are followed by the exchange
    void finish_halo_exchange() {}

    // Beyond the interface: used by this project's Solver.
    void set_potential(int which, const double *v);   ///< raw V of one component, for energy()
    void energy(double out[12]) const; ///< the two-component energies (tsmi_pair_energy)
    tsmi_pair *pair() const { return pair_; }

private:
    void flush() const;                ///< issues queued work that no rabi_coupling followed
    tsmi_pair *pair_;
    int precision_;
    int width_, height_;
    bool imag_;
    double delta_t_;
    int state_index_;
    mutable int queued_;
    mutable bool exchange_;
    mutable bool rot_pending_;
    mutable double rot_var_;
};

#endif
