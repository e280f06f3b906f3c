// tsmi_pair.hip -- C ABI of the two-component step (include/tsmi_hip.h, tsmi_pair_*).
//
// A two-component system (Hamiltonian2Component with the two-state Solver and
// CPUBlock, src/trottersuzuki.h:459-497, src/solver.cpp:43-60,150-242,
// src/cpukernel.cpp:274-392,551-667) is held as two single-component contexts on
// one stream -- one per component, each with its own mass, self coupling, E
// table and raw potential, their ping-pong buffers flipped together -- plus the
// Rabi and inter-species constants and the imaginary-time norm targets, which
// live on the device because the joint renormalisation changes them every step.
// A step of both components is one pair_kernel launch (pair_kernel.hpp); in FP32
// real time on closed or periodic lattices tsmi_pair_evolve runs two steps per
// launch (pair2_kernel) on the components' 8-wide halo layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "context.hpp"
#include "pair_kernel.hpp"
#include "tsmi_hip.h"

using namespace tsmi;

namespace {

constexpr int kBlock = 256;
constexpr int kVecPair32 = 2;   // FP32 columns per lane of pair_kernel and pair2_kernel (FP64: 2)
static_assert(kBlock == kPair2Block, "pair2_kernel's workgroup size is its launch bound");
constexpr int kRabiBlocksMax = 1024;

#define TSMI_HIPP(call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return set_error(TSMI_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct tsmi_pair {
    tsmi_ctx *c[2] = {nullptr, nullptr};
    int esz = 0, vec = 0;
    bool imag = false;
    double dt = 0, c_ab = 0, omega_r = 0, omega_i = 0, g_ab = 0, tot_norm = 0;
    int W = 0, H = 0, pitch = 0;
    int n_strips = 0, seg_rows = 0, n_segs = 0, n_waves = 0, rabi_blocks = 0;
    // two steps per pass (pair2_kernel) over the whole allocation, Wd x Hd
    bool fused = false, halo8 = false;
    int Wd = 0, Hd = 0, n_strips2 = 0, seg_rows2 = 0, n_segs2 = 0, n_waves2 = 0;
    int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;
    // device scalars: divisor pairs [2][2] (a pending pair is read while the
    // next one is written), norm targets [2]
    double *dv = nullptr;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    int slot = 0;
    bool pending = false;
};

namespace {

void free_pair(tsmi_pair *p) {
    if (p->c[1]) tsmi_destroy(p->c[1]);   // runs on c[0]'s stream: released first
    if (p->c[0]) tsmi_destroy(p->c[0]);
    void *ptrs[] = {p->dv, p->partials, p->counter};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

bool same_lattice(const tsmi_desc &a, const tsmi_desc &b) {
    return a.dim_x == b.dim_x && a.dim_y == b.dim_y && a.periodic_x == b.periodic_x &&
           a.periodic_y == b.periodic_y && a.halo_x == b.halo_x && a.halo_y == b.halo_y &&
           a.imag_time == b.imag_time && a.precision == b.precision && a.delta_t == b.delta_t &&
           a.delta_x == b.delta_x && a.delta_y == b.delta_y && a.length_x == b.length_x &&
           a.length_y == b.length_y && a.tile_rows == b.tile_rows;
}

// The rotation constants of CPUBlock::rabi_coupling(var, dt) (src/cpukernel.cpp:581-608).
void rabi_coefs(const tsmi_pair *p, double var, double *cc, double *cs_r, double *cs_i) {
    const double cr = 0.5 * p->omega_r, ci = 0.5 * p->omega_i;
    const double nrm = std::sqrt(cr * cr + ci * ci);
    const double ang = -p->dt * var * nrm;
    if (p->imag) {
        *cc = std::cosh(ang);
        *cs_r = nrm == 0 ? 0 : cr / nrm * std::sinh(ang);
        *cs_i = nrm == 0 ? 0 : ci / nrm * std::sinh(ang);
    } else {
        *cc = std::cos(ang);
        *cs_r = nrm == 0 ? 0 : cr / nrm * std::sin(ang);
        *cs_i = nrm == 0 ? 0 : ci / nrm * std::sin(ang);
    }
}

bool rabi_on(const tsmi_pair *p) { return p->omega_r != 0.0 || p->omega_i != 0.0; }

double *div_pair(tsmi_pair *p, int s) { return p->dv + 2 * s; }
double *norms(tsmi_pair *p) { return p->dv + 4; }

// One step of both components; in real time the Rabi rotation by var leaves the
// pipelines with the rows when rotate.
template <typename T, bool IMAG, bool PH>
int launch_pair_step(tsmi_pair *p, double var, bool rotate) {
    PairParams<T> q;
    std::memset(&q, 0, sizeof(q));
    const int cur = p->c[0]->cur, nb = 1 - cur;
    for (int i = 0; i < 2; ++i) {
        tsmi_ctx *c = p->c[i];
        q.in_re[i] = (const T *)c->psi_re[cur];
        q.in_im[i] = (const T *)c->psi_im[cur];
        q.out_re[i] = (T *)c->psi_re[nb];
        q.out_im[i] = (T *)c->psi_im[nb];
        q.e_re[i] = (const T *)c->e_re;
        q.e_im[i] = (const T *)c->e_im;
        q.ph[i] = (const T *)c->ph;
        q.mag[i] = IMAG ? 1.0 : c->corr_abs;
        q.mH[i] = (T)c->mH; q.tH[i] = (T)c->tH; q.mV[i] = (T)c->mV; q.tV[i] = (T)c->tV;
        q.unH[i] = (T)c->unH; q.unV[i] = (T)c->unV;
        q.c_self[i] = (T)c->c_nl_s;
    }
    q.W = p->W; q.H = p->H; q.pitch = p->pitch;
    q.n_strips = p->n_strips; q.n_waves = p->n_waves; q.seg_rows = p->seg_rows;
    q.c_ab = (T)p->c_ab;
    if (!IMAG && rotate) {
        double cc, sr, si;
        rabi_coefs(p, var, &cc, &sr, &si);
        q.cc = cc; q.cs_r = sr; q.cs_i = si;
        q.rabi = 1;
    }
    if (IMAG) {
        if (p->pending) {
            q.div_in[0] = div_pair(p, p->slot);
            q.div_in[1] = div_pair(p, p->slot) + 1;
        }
        q.div_out = div_pair(p, 1 - p->slot);
        q.partials = p->partials;
        q.counter = p->counter;
        q.norm = norms(p);
        q.dxdy = p->c[0]->d.delta_x * p->c[0]->d.delta_y;
        q.ix0 = p->ix0; q.ix1 = p->ix1; q.iy0 = p->iy0; q.iy1 = p->iy1;
    }
    constexpr int VEC = sizeof(T) == 4 ? kVecPair32 : 2;
    const int blocks = (p->n_waves + kBlock / kWave - 1) / (kBlock / kWave);
    // FP64 real time: the fast beat needs one scaled form for both components on both axes
    const int mode = p->c[0]->mode == p->c[1]->mode ? p->c[0]->mode : -1;
    if constexpr (IMAG || sizeof(T) == 4)
        hipLaunchKernelGGL((pair_kernel<T, IMAG, VEC, PH, -1>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    else if (mode == 0)
        hipLaunchKernelGGL((pair_kernel<T, IMAG, VEC, PH, 0>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    else if (mode == 1)
        hipLaunchKernelGGL((pair_kernel<T, IMAG, VEC, PH, 1>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    else
        hipLaunchKernelGGL((pair_kernel<T, IMAG, VEC, PH, -1>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    TSMI_HIPP(hipGetLastError());
    p->c[0]->cur = nb;
    p->c[1]->cur = nb;
    if (IMAG) {
        p->slot = 1 - p->slot;
        p->pending = true;
    }
    return TSMI_OK;
}

// Rabi rotation by var (rotate) and/or the joint renormalisation (normalize) of
// both current tiles; with neither, only the pending divisors are applied.
template <typename T>
int launch_rabi(tsmi_pair *p, double var, bool rotate, bool normalize) {
    RabiParams<T> q;
    std::memset(&q, 0, sizeof(q));
    const int cur = p->c[0]->cur;
    for (int i = 0; i < 2; ++i) {
        q.re[i] = (T *)p->c[i]->psi_re[cur];
        q.im[i] = (T *)p->c[i]->psi_im[cur];
    }
    q.W = p->W; q.H = p->H; q.pitch = p->pitch;
    if (rotate) {
        double cc, sr, si;
        rabi_coefs(p, var, &cc, &sr, &si);
        q.cc = cc; q.cs_r = sr; q.cs_i = si;
        q.rotate = 1;
    }
    if (p->pending) {
        q.div_in[0] = div_pair(p, p->slot);
        q.div_in[1] = div_pair(p, p->slot) + 1;
    }
    q.normalize = normalize ? 1 : 0;
    if (normalize) {
        q.partials = p->partials;
        q.counter = p->counter;
        q.div_out = div_pair(p, 1 - p->slot);
        q.norm = norms(p);
        q.tot_norm = p->tot_norm;
        q.dxdy = p->c[0]->d.delta_x * p->c[0]->d.delta_y;
        q.ix0 = p->ix0; q.ix1 = p->ix1; q.iy0 = p->iy0; q.iy1 = p->iy1;
    }
    hipStream_t s = p->c[0]->stream;
    if (p->imag)
        hipLaunchKernelGGL((rabi_kernel<T, true>), dim3(p->rabi_blocks), dim3(kRabiThreads), 0, s, q);
    else
        hipLaunchKernelGGL((rabi_kernel<T, false>), dim3(p->rabi_blocks), dim3(kRabiThreads), 0, s, q);
    TSMI_HIPP(hipGetLastError());
    if (normalize) {
        p->slot = 1 - p->slot;
        p->pending = true;
    } else {
        p->pending = false;   // the pending divisors are now applied
    }
    return TSMI_OK;
}

int rabi_pass(tsmi_pair *p, double var, bool rotate, bool normalize) {
    return p->esz == 4 ? launch_rabi<float>(p, var, rotate, normalize) : launch_rabi<double>(p, var, rotate, normalize);
}

template <typename T, bool PH>
int pair_step_form(tsmi_pair *p, double var, bool rotate) {
    return p->imag ? launch_pair_step<T, true, PH>(p, var, rotate) : launch_pair_step<T, false, PH>(p, var, rotate);
}

// the one-plane form of E when both components have it
int pair_step(tsmi_pair *p, double var, bool rotate) {
    const bool ph = p->c[0]->ph_mode && p->c[1]->ph_mode;
    if (p->esz == 4) return ph ? pair_step_form<float, true>(p, var, rotate) : pair_step_form<float, false>(p, var, rotate);
    return ph ? pair_step_form<double, true>(p, var, rotate) : pair_step_form<double, false>(p, var, rotate);
}

// the periodic exchange of both components' current tiles (on the 8-wide halo
// layout the full wrap, whose inner 4 cells are the reference's exchange)
int exchange_both(tsmi_pair *p) {
    for (int c = 0; c < 2; ++c) {
        tsmi_ctx *x = p->c[c];
        if (int st = ctx_exchange(x, x->psi_base[x->cur][0], x->psi_base[x->cur][1])) return st;
    }
    return TSMI_OK;
}

// Two real-time FP32 steps of both components in one pass over the allocations;
// the Rabi rotation by var1 ends the first, by var2 the second.  Phase form of
// E only: the E-form instance needs more than 256 VGPRs (it spills).
int pair_step2(tsmi_pair *p, double var1, double var2, bool rotate) {
    constexpr bool PH = true;
    using T = float;
    PairParams<T> q;
    std::memset(&q, 0, sizeof(q));
    const int cur = p->c[0]->cur, nb = 1 - cur;
    for (int i = 0; i < 2; ++i) {
        tsmi_ctx *c = p->c[i];
        q.in_re[i] = (const T *)c->psi_base[cur][0];
        q.in_im[i] = (const T *)c->psi_base[cur][1];
        q.out_re[i] = (T *)c->psi_base[nb][0];
        q.out_im[i] = (T *)c->psi_base[nb][1];
        q.e_re[i] = (const T *)(c->halo8 ? c->e8[0] : c->e_base[0]);
        q.e_im[i] = (const T *)(c->halo8 ? c->e8[1] : c->e_base[1]);
        q.ph[i] = (const T *)(c->halo8 ? c->ph8 : c->ph_base);
        q.mag[i] = c->corr_abs;
        q.mH[i] = (T)c->mH; q.tH[i] = (T)c->tH; q.mV[i] = (T)c->mV; q.tV[i] = (T)c->tV;
        q.unH[i] = (T)c->unH; q.unV[i] = (T)c->unV;
        q.c_self[i] = (T)c->c_nl_s;
    }
    q.W = p->Wd; q.H = p->Hd; q.pitch = p->pitch;
    q.n_strips = p->n_strips2; q.n_waves = p->n_waves2; q.seg_rows = p->seg_rows2;
    q.c_ab = (T)p->c_ab;
    if (rotate) {
        rabi_coefs(p, var1, &q.cc, &q.cs_r, &q.cs_i);
        rabi_coefs(p, var2, &q.cc2, &q.cs_r2, &q.cs_i2);
        q.rabi = 1;
    }
    const int blocks = (p->n_waves2 + kBlock / kWave - 1) / (kBlock / kWave);
    // the fast beat needs one scaled form for both components on both axes
    const int mode = p->c[0]->mode == p->c[1]->mode ? p->c[0]->mode : -1;
    if (mode == 0)
        hipLaunchKernelGGL((pair2_kernel<T, kVecPair32, PH, 0>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    else if (mode == 1)
        hipLaunchKernelGGL((pair2_kernel<T, kVecPair32, PH, 1>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    else
        hipLaunchKernelGGL((pair2_kernel<T, kVecPair32, PH, -1>), dim3(blocks), dim3(kBlock), 0, p->c[0]->stream, q);
    TSMI_HIPP(hipGetLastError());
    p->c[0]->cur = nb;
    p->c[1]->cur = nb;
    return TSMI_OK;
}

// Solver::evolve(n) with two steps per pass (src/solver.cpp:186-230), scheduled
// as the single-component evolve_steps does: on closed axes the pairs replace
// single steps from the first; on periodic axes the first step runs alone on
// the reference tile, every pass starts from a freshly wrapped 8-wide halo, and
// the last step runs alone, leaving the reference's unrefreshed halo.  The last
// step's rotation is by 1/2 (src/solver.cpp:222-225).
int evolve_fused(tsmi_pair *p, int n, bool rabi) {
    auto var_of = [n](int i) { return i == n - 1 ? 0.5 : 1.0; };
    int st = TSMI_OK, i = 0;
    if (!p->halo8) {
        for (; i + 1 < n && !st; i += 2) st = pair_step2(p, var_of(i), var_of(i + 1), rabi);
        if (!st && i < n) st = pair_step(p, var_of(i), rabi);
        return st;
    }
    if (n >= 3) {
        st = pair_step(p, 1.0, rabi);
        if (!st) st = exchange_both(p);
        for (i = 1; n - i >= 3 && !st; i += 2) {
            st = pair_step2(p, 1.0, 1.0, rabi);
            if (!st) st = exchange_both(p);
        }
    }
    for (; i < n && !st; ++i) {
        st = pair_step(p, var_of(i), rabi);
        if (!st && i != n - 1) st = exchange_both(p);
    }
    return st;
}

// Divides both current tiles by their pending divisors, if any.
int settle(tsmi_pair *p) { return p->pending ? rabi_pass(p, 0.0, false, false) : TSMI_OK; }

int check_ready(tsmi_pair *p, const char *who) {
    if (!p) return set_error(TSMI_E_INVALID, std::string(who) + ": null pair");
    for (int i = 0; i < 2; ++i)
        if (!p->c[i]->has_state || !p->c[i]->has_e)
            return set_error(TSMI_E_INVALID, std::string(who) + ": both states and E tables must be set");
    return TSMI_OK;
}

template <typename T, bool IMAG>
hipError_t pair_occupancy(int *blocks) {
    constexpr int VEC = sizeof(T) == 4 ? kVecPair32 : 2;
    constexpr int MODE = (IMAG || sizeof(T) == 4) ? -1 : 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, pair_kernel<T, IMAG, VEC, true, MODE>, kBlock, 0);
}

hipError_t occupancy_of(int esz, bool imag, int *blocks) {
    if (esz == 4) return imag ? pair_occupancy<float, true>(blocks) : pair_occupancy<float, false>(blocks);
    return imag ? pair_occupancy<double, true>(blocks) : pair_occupancy<double, false>(blocks);
}

tsmi_ctx *component(tsmi_pair *p, int which, const char *who) {
    if (!p || which < 0 || which > 1) {
        set_error(TSMI_E_INVALID, std::string(who) + ": need a pair and which = 0 or 1");
        return nullptr;
    }
    return p->c[which];
}

}  // namespace

extern "C" {

int tsmi_pair_create(const tsmi_pair_desc *pd, tsmi_pair **out) {
    if (!pd || !out) return set_error(TSMI_E_INVALID, "tsmi_pair_create: null argument");
    *out = nullptr;
    if (!same_lattice(pd->a, pd->b))
        return set_error(TSMI_E_INVALID, "tsmi_pair_create: both components need the same lattice, time step, "
                                         "time direction, precision and tile_rows");
    if (pd->a.coupling_lhy != 0.0 || pd->b.coupling_lhy != 0.0)
        return set_error(TSMI_E_UNSUPPORTED, "tsmi_pair_create: a two-component system has no LHY term");
    // two steps per pass: FP32 real time, both axes closed or both periodic
    const tsmi_desc &da = pd->a;
    const bool fusable = da.precision == 32 && !da.imag_time &&
                         ((da.periodic_x && da.periodic_y) || (!da.periodic_x && !da.periodic_y));
    if (da.sweeps < 0 || da.sweeps > 2) return set_error(TSMI_E_INVALID, "tsmi_pair_create: sweeps must be 0, 1 or 2");
    if (da.sweeps == 2 && !fusable)
        return set_error(TSMI_E_UNSUPPORTED, "tsmi_pair_create: two steps per pass need FP32, real time and both "
                                             "axes closed or both periodic");
    const bool fused = fusable && da.sweeps != 1;
    tsmi_pair *p = new tsmi_pair();
    auto bail = [&](int code) {
        free_pair(p);
        delete p;
        return code;
    };
    for (int i = 0; i < 2; ++i) {
        tsmi_desc d = i == 0 ? pd->a : pd->b;
        d.sweeps = fused ? 2 : 1;   // 2: the 8-wide halo layout on periodic axes, E wrapped over it
        if (int st = tsmi_create(&d, &p->c[i])) return bail(st);
    }
    if (int st = tsmi_set_stream(p->c[1], p->c[0]->stream)) return bail(st);
    const tsmi_desc &d = pd->a;
    p->esz = d.precision == 32 ? 4 : 8;
    p->vec = p->esz == 4 ? kVecPair32 : 2;
    p->imag = d.imag_time != 0;
    p->dt = d.delta_t;
    p->g_ab = pd->coupling_ab;
    p->c_ab = pd->coupling_ab * d.delta_t;   // src/cpukernel.cpp:329
    p->omega_r = pd->omega_r;
    p->omega_i = pd->omega_i;
    tsmi_ctx *c0 = p->c[0];
    p->W = c0->W; p->H = c0->H; p->pitch = c0->pitch;
    p->ix0 = c0->ix0; p->ix1 = c0->ix1; p->iy0 = c0->iy0; p->iy1 = c0->iy1;
    // strips of kWave*vec - 8 output columns; segments sized to one round of
    // resident waves of pair_kernel, at least 24 rows
    const int sw = kWave * p->vec - 2 * kApron;
    p->n_strips = (p->W + sw - 1) / sw;
    int seg = d.tile_rows;
    if (seg <= 0) {
        seg = std::max(24, (p->H + 1) / 2);
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            occupancy_of(p->esz, p->imag, &per_cu) == hipSuccess && cus > 0 && per_cu > 0) {
            const long resident = (long)cus * per_cu * (kBlock / kWave);
            const long segs = std::max<long>(1, resident / p->n_strips);
            seg = (int)std::max<long>(24, (p->H + segs - 1) / segs);
        }
    }
    seg = std::min(seg, max_seg_rows(p->pitch, 2 * kApron));
    seg = (seg + 1) / 2 * 2;   // segments start on even rows
    p->seg_rows = std::min(seg, p->H + (p->H & 1));
    p->n_segs = (p->H + p->seg_rows - 1) / p->seg_rows;
    p->n_waves = p->n_strips * p->n_segs;
    p->rabi_blocks = (int)std::min<long>(kRabiBlocksMax, ((long)p->W * p->H + kRabiThreads - 1) / kRabiThreads);
    p->fused = fused;
    p->halo8 = c0->halo8;
    p->Wd = c0->Wd; p->Hd = c0->Hd;
    if (fused) {
        // strips of kWave*vec - 16 output columns over the allocation; segments
        // sized to one round of resident waves of pair2_kernel, at least 16 rows
        const int sw2 = kWave * kVecPair32 - 4 * kApron;
        p->n_strips2 = (p->Wd + sw2 - 1) / sw2;
        int seg2 = d.tile_rows;
        if (seg2 <= 0) {
            seg2 = std::max(16, (p->Hd + 1) / 2);
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pair2_kernel<float, kVecPair32, true, 0>, kBlock,
                                                             0) == hipSuccess &&
                cus > 0 && per_cu > 0) {
                const long resident = (long)cus * per_cu * (kBlock / kWave);
                const long segs = std::max<long>(1, resident / p->n_strips2);
                seg2 = (int)std::max<long>(16, (p->Hd + segs - 1) / segs);
            }
        }
        seg2 = std::min(seg2, max_seg_rows(p->pitch, 4 * kApron));
        seg2 = (seg2 + 1) / 2 * 2;   // segments start on even rows
        p->seg_rows2 = std::min(seg2, p->Hd + (p->Hd & 1));
        p->n_segs2 = (p->Hd + p->seg_rows2 - 1) / p->seg_rows2;
        p->n_waves2 = p->n_strips2 * p->n_segs2;
    }
    // norm targets at kernel construction (src/solver.cpp:153-159, src/cpukernel.cpp:321-325)
    const double nt[6] = {1.0, 1.0, 1.0, 1.0, pd->a.norm_target, pd->b.norm_target};
    p->tot_norm = nt[4] + nt[5];
    const size_t np = 2 * (size_t)std::max(p->n_waves, p->rabi_blocks);
    hipError_t e = hipMalloc((void **)&p->dv, sizeof(nt));
    if (e == hipSuccess) e = hipMalloc((void **)&p->partials, np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&p->counter, 64);
    if (e != hipSuccess) return bail(set_error(TSMI_E_OOM, std::string("tsmi_pair_create: ") + hipGetErrorString(e)));
    hipStream_t s = c0->stream;
    e = hipMemcpyAsync(p->dv, nt, sizeof(nt), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(p->counter, 0, 64, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("tsmi_pair_create: ") + hipGetErrorString(e)));
    *out = p;
    return TSMI_OK;
}

void tsmi_pair_destroy(tsmi_pair *p) {
    if (!p) return;
    if (p->c[0]) (void)hipStreamSynchronize(p->c[0]->stream);
    free_pair(p);
    delete p;
}

int tsmi_pair_set_state(tsmi_pair *p, int32_t which, const double *re, const double *im) {
    tsmi_ctx *c = component(p, which, "tsmi_pair_set_state");
    if (!c) return TSMI_E_INVALID;
    return tsmi_set_state(c, re, im);
}

int tsmi_pair_set_exp_potential(tsmi_pair *p, int32_t which, const double *e_re, const double *e_im) {
    tsmi_ctx *c = component(p, which, "tsmi_pair_set_exp_potential");
    if (!c) return TSMI_E_INVALID;
    return tsmi_set_exp_potential(c, e_re, e_im);
}

int tsmi_pair_set_potential(tsmi_pair *p, int32_t which, const double *v) {
    tsmi_ctx *c = component(p, which, "tsmi_pair_set_potential");
    if (!c) return TSMI_E_INVALID;
    return tsmi_set_potential(c, v);
}

int tsmi_pair_evolve(tsmi_pair *p, int32_t n_steps, int32_t initial_half) {
    if (int st = check_ready(p, "tsmi_pair_evolve")) return st;
    if (n_steps < 0) return set_error(TSMI_E_INVALID, "tsmi_pair_evolve: negative step count");
    const bool rabi = rabi_on(p);
    int st = TSMI_OK;
    // src/solver.cpp:173-177
    if (initial_half && rabi) st = rabi_pass(p, 0.5, true, false);
    // two steps per pass where both E tables have the phase form
    if (p->fused && p->c[0]->ph_mode && p->c[1]->ph_mode) return st ? st : evolve_fused(p, n_steps, rabi);
    for (int i = 0; i < n_steps && !st; ++i) {
        const bool last = i == n_steps - 1;
        const double var = last ? 0.5 : 1.0;   // src/solver.cpp:222-225
        st = pair_step(p, var, rabi);
        // the periodic exchange of both components (src/solver.cpp:203-219);
        // the Rabi rotation and the renormalisations act cell by cell and
        // commute with it
        if (!st && !last) st = exchange_both(p);
        if (!st && p->imag && rabi) st = rabi_pass(p, var, true, true);   // src/solver.cpp:225-226
    }
    // the state at the end of a call is the renormalised one (src/cpukernel.cpp:560-563, :650-664)
    if (!st) st = settle(p);
    return st;
}

int tsmi_pair_step(tsmi_pair *p, int32_t exchange, double rabi_var) {
    if (int st = check_ready(p, "tsmi_pair_step")) return st;
    int st = pair_step(p, rabi_var, !p->imag && rabi_on(p) && rabi_var != 0.0);
    if (!st && exchange) st = exchange_both(p);
    return st;
}

int tsmi_pair_rabi(tsmi_pair *p, double var, int32_t normalize) {
    if (int st = check_ready(p, "tsmi_pair_rabi")) return st;
    const bool rot = rabi_on(p) && var != 0.0;
    const bool nrm = normalize && p->imag && rabi_on(p);
    if (!rot && !nrm) return TSMI_OK;
    return rabi_pass(p, var, rot, nrm);
}

int tsmi_pair_get_state(tsmi_pair *p, int32_t which, double *re, double *im) {
    tsmi_ctx *c = component(p, which, "tsmi_pair_get_state");
    if (!c) return TSMI_E_INVALID;
    if (!re || !im) return set_error(TSMI_E_INVALID, "tsmi_pair_get_state: null argument");
    if (int st = settle(p)) return st;
    return tsmi_get_sample(c, (size_t)c->W, 0, 0, (size_t)c->W, (size_t)c->H, re, im);
}

int tsmi_pair_energy(tsmi_pair *p, double out[12]) {
    if (!p || !out) return set_error(TSMI_E_INVALID, "tsmi_pair_energy: null argument");
    double ea[6], eb[6];
    int st = settle(p);
    if (!st) st = tsmi_energy(p->c[0], ea);
    if (!st) st = tsmi_energy(p->c[1], eb);
    if (st) return st;
    tsmi_ctx *a = p->c[0], *b = p->c[1];
    const int blocks = std::max(1, std::min(a->obs_blocks, (p->iy1 - p->iy0)));
    const int cur = a->cur;
    if (p->esz == 4)
        hipLaunchKernelGGL((pair_cross_kernel<float>), dim3(blocks), dim3(kRabiThreads), 0, a->stream,
                           (const float *)a->psi_re[cur], (const float *)a->psi_im[cur], (const float *)b->psi_re[cur],
                           (const float *)b->psi_im[cur], p->pitch, p->ix0, p->ix1, p->iy0, p->iy1, p->omega_r,
                           p->omega_i, a->obs_partials);
    else
        hipLaunchKernelGGL((pair_cross_kernel<double>), dim3(blocks), dim3(kRabiThreads), 0, a->stream,
                           (const double *)a->psi_re[cur], (const double *)a->psi_im[cur],
                           (const double *)b->psi_re[cur], (const double *)b->psi_im[cur], p->pitch, p->ix0, p->ix1,
                           p->iy0, p->iy1, p->omega_r, p->omega_i, a->obs_partials);
    TSMI_HIPP(hipGetLastError());
    std::vector<double> h((size_t)blocks * 4);
    TSMI_HIPP(hipMemcpyAsync(h.data(), a->obs_partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    TSMI_HIPP(hipStreamSynchronize(a->stream));
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (int i = 0; i < blocks; ++i)
        for (int j = 0; j < 4; ++j) {
            const double v = h[(size_t)i * 4 + j], t = s[j] + v;
            c[j] += (std::fabs(s[j]) >= std::fabs(v)) ? (s[j] - t) + v : (v - t) + s[j];
            s[j] = t;
        }
    const double na = s[0] + c[0], nb = s[1] + c[1], s4 = s[2] + c[2], sr = s[3] + c[3];
    // src/solver.cpp:546-554
    const double inter = p->g_ab * s4 / (na * nb);
    const double rabi = 0.5 * sr / (na * nb);
    out[1] = ea[1]; out[2] = eb[1];
    out[3] = ea[2]; out[4] = eb[2];
    out[5] = ea[3]; out[6] = eb[3];
    out[7] = inter; out[8] = rabi;
    out[0] = ea[1] + ea[2] + ea[3] + eb[1] + eb[2] + eb[3] + inter + rabi;
    out[9] = ea[5]; out[10] = eb[5];
    out[11] = ea[5] + eb[5];
    return TSMI_OK;
}

int tsmi_pair_geometry(const tsmi_pair *p, int32_t out[6]) {
    if (!p || !out) return set_error(TSMI_E_INVALID, "tsmi_pair_geometry: null argument");
    out[0] = p->W; out[1] = p->H; out[2] = p->pitch;
    out[3] = p->n_strips; out[4] = p->n_segs; out[5] = p->seg_rows;
    return TSMI_OK;
}

int tsmi_pair_synchronize(tsmi_pair *p) {
    if (!p) return set_error(TSMI_E_INVALID, "tsmi_pair_synchronize: null pair");
    TSMI_HIPP(hipStreamSynchronize(p->c[0]->stream));
    return TSMI_OK;
}

}  // extern "C"
