// context.hpp -- the single-component context (tsmi_ctx) shared by the
// library's translation units.
//
// A context owns two device copies of the tile (ping-pong: a step reads one and
// writes the other, as CPUBlock does with its two buffers,
// src/cpukernel.cpp:244-247), the E planes, the raw potential and the
// coordinate tables.  The two-component context (tsmi_pair.hip) is built from
// two of them, one per component, on one stream.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "tsmi_hip.h"

struct tsmi_ctx {
    tsmi_desc d;
    int W = 0, H = 0, pitch = 0, esz = 0, vec = 0;
    int n_strips = 0, n_segs = 0, seg_rows = 0, n_waves = 0;
    // Views at the origin of the reference tile (W x H, row pitch `pitch`).  With
    // the 8-wide halo layout they sit 4 rows and 4 columns inside allocations of
    // Wd x Hd = (W + 8) x (H + 8); otherwise view and allocation coincide.
    void *psi_re[2] = {nullptr, nullptr}, *psi_im[2] = {nullptr, nullptr};
    void *e_re = nullptr, *e_im = nullptr;
    void *psi_base[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    void *e_base[2] = {nullptr, nullptr};
    void *e8[2] = {nullptr, nullptr};     // E wrapped over the whole 8-wide halo (fused passes)
    // E corr as one plane (phase in real time, exponent in imaginary time): the
    // steps read it instead of E when ph_mode (every |E| = 1, or every E > 0)
    void *ph_base = nullptr, *ph = nullptr, *ph8 = nullptr;
    bool ph_mode = false;
    unsigned *flag = nullptr;             // set by the E import when E has no phase form
    double *v_base = nullptr;
    int Wd = 0, Hd = 0;
    size_t off_el = 0;                    // elements from allocation to view origin
    bool halo8 = false;
    bool nt_out = false;                  // FP64 real time: output rows stored non-temporal
    // kernel timing (tsmi_kernel_timing): one event pair per step-kernel launch
    bool timing = false;
    std::vector<hipEvent_t> ev[2][2];     // [fused?][start/end]
    size_t ev_used[2] = {0, 0};
    // fused two-step passes (fused_kernel.hpp)
    int sweeps = 1;
    int n_strips2 = 0, n_segs2 = 0, seg_rows2 = 0, n_waves2 = 0;
    int geo2_key = -1;                    // (form, E form) the segments were sized for; -1: none
    double *v = nullptr, *xc = nullptr, *yc = nullptr;
    // imaginary time: two slots of per-wave partial sums (one per step parity),
    // their counts, and one divisor for host-side readers (divisor_kernel)
    double *div = nullptr;
    double *partials = nullptr;
    int *part_n = nullptr;
    double *obs_partials = nullptr;      // kObsBlocksMax * OBS_COUNT block partials
    double *obs_out = nullptr;           // kObsOut results of the last observables pass
    int obs_blocks = 0;
    int cur = 0, div_slot = 0;
    bool div_pending = false, has_state = false, has_e = false, has_v = false;
    hipStream_t own = nullptr, stream = nullptr;
    double aH = 0, bH = 0, aV = 0, bV = 0, c_nl = 0, c_lhy = 0;   // FP64, as the reference
    // scaled-frame factors actually used by the kernel (values of the kernel precision)
    double mH = 1, tH = 0, mV = 1, tV = 0, unH = 1, unV = 1;
    double corr_re = 1.0, corr_im = 0.0;   // folded into E
    double c_nl_s = 0, c_lhy_s = 0;        // couplings rescaled to the mid-step frame
    int mode = 0;                          // 0 / 1: both axes in form A / B; -1: mixed
    int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;
    // reverse sweep (real time, closed axes, no nonlinearity)
    bool adj_ok = false;
    double amH = 1, atH = 0, amV = 1, atV = 0, aunH = 1, aunV = 1;   // frame of the step at -dt
    double corr_abs = 1.0;                                         // |corr| = |kappa|^2
    void *ck = nullptr;               // checkpoints: 2 planes each
    size_t ck_cap = 0;
    int ck_every = 0, ck_steps = 0, ck_count = 0;
    bool ck_valid = false;
    void *adj = nullptr;              // 8 planes: adjoint re/im x2, reversed state re/im x2
    double *grad = nullptr;           // FP64, pitched
};

namespace tsmi {

// The periodic halo refresh of one state allocation of c (re/im planes in the
// context precision), asynchronous on c's stream.
int ctx_exchange(tsmi_ctx *c, void *re, void *im);

}  // namespace tsmi
