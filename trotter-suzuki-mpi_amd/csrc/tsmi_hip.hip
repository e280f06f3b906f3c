// tsmi_hip.hip -- C ABI of the gfx950 Trotter-Suzuki step (include/tsmi_hip.h).
//
// A context owns two device copies of the tile (ping-pong: a step reads one and
// writes the other, as CPUBlock does with its two buffers, src/cpukernel.cpp:244-247),
// the E planes, the raw potential and the coordinate tables, all in the context
// precision except the FP64 potential/coordinates used by the reductions.
// Rows are padded to a multiple of 64 elements; the padding is zero and never
// written, so lanes that read past the tile edge see finite values.
#include "tsmi_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "adjoint_kernel.hpp"
#include "common.hpp"
#include "context.hpp"
#include "frame.hpp"
#include "fused_kernel.hpp"
#include "observables.hpp"
#include "step_kernel.hpp"

using namespace tsmi;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

// The device array behind tsmi_energy_device_noalloc: block partials of one
// observables pass (kObsBlocksMax x OBS_COUNT doubles).  Each GPU holds its own
// copy; host code reaches it only through the address hipGetSymbolAddress
// returns on that device.
static __device__ __align__(256) double scratch[tsmi::kObsBlocksMax * tsmi::OBS_COUNT];
static_assert(sizeof(scratch) == TSMI_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");
static_assert(TSMI_NOALLOC_MAX_COLUMNS == tsmi::kObsBlocksMax * tsmi::kObsBlockCols, "the header states the limit");

namespace tsmi {

int set_error(int code, const std::string &msg) { return fail(code, msg); }

int check_desc(const tsmi_desc &d, const char *who) {
    const std::string w(who);
    if (d.dim_x < 1 || d.dim_y < 1) return fail(TSMI_E_INVALID, w + ": lattice dimensions must be positive");
    if (d.precision != 32 && d.precision != 64) return fail(TSMI_E_INVALID, w + ": precision must be 32 or 64");
    // a Lattice1D (one row, closed y) has no y halo (src/model.cpp:68-69)
    const bool chain = d.dim_y == 1 && !d.periodic_y && d.halo_y == 0;
    if (d.halo_x != 4 || (d.halo_y != 4 && !chain))
        return fail(TSMI_E_UNSUPPORTED, w + ": only the halo width 4 of a non-rotating lattice is supported");
    if (!(d.delta_x > 0) || !(d.delta_y > 0) || !(d.mass != 0))
        return fail(TSMI_E_INVALID, w + ": lattice spacing and mass must be non-zero");
    if ((d.periodic_x && d.dim_x < 2 * d.halo_x) || (d.periodic_y && d.dim_y < 2 * d.halo_y))
        return fail(TSMI_E_INVALID, w + ": a periodic axis needs at least 2*halo points");
    if (d.tile_rows < 0) return fail(TSMI_E_INVALID, w + ": tile_rows must be >= 0");
    // the widest allocation: the 8-wide halo layout of two steps per pass
    const long long W = (long long)d.dim_x + (d.periodic_x ? 4 * d.halo_x : 0);
    const long long H = (long long)d.dim_y + (d.periodic_y ? 4 * d.halo_y : 0);
    const long long pitch = (W + 63) / 64 * 64;
    // a wave addresses its segment with 32-bit byte offsets: the smallest
    // segment of a two-step pass (2 rows and 16 apron rows) of FP64 must stay
    // below 2^31 bytes
    if (pitch * 8 * (2 + 4 * kApron) >= (1LL << 31) || H > (1 << 30))
        return fail(TSMI_E_INVALID, w + ": lattice too large");
    return TSMI_OK;
}

int max_seg_rows(long long pitch, int apron_rows) {
    const long long rows = ((1LL << 31) - 1) / (pitch * 8) - apron_rows;
    return (int)std::max<long long>(2, std::min<long long>(rows, 1 << 30) / 2 * 2);
}

int plan_seg_rows(int H, int strips, long resident, long simds, int floor_rows, bool small12, int tile_rows,
                  long long pitch, int apron_rows) {
    int seg = tile_rows;
    if (seg <= 0) {
        seg = std::max<int>(floor_rows, (H + 1) / 2);
        if (resident > 0) {
            const long segs = std::max<long>(1, resident / std::max(1, strips));
            seg = (int)std::max<long>(floor_rows, (H + segs - 1) / segs);
            if (small12 && (long)strips * ((H + seg - 1) / seg) < simds)
                seg = (int)std::max<long>(12, (H + segs - 1) / segs);
        }
    }
    seg = std::min(seg, max_seg_rows(pitch, apron_rows));
    seg = (seg + 1) / 2 * 2;
    return std::min(seg, H);
}

}  // namespace tsmi

namespace {

#define TSMI_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(TSMI_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;          // 4 wavefronts per workgroup

// FP32 columns per lane of the single-step and of the fused two-step kernel
// (FP64 uses 2: one 16-byte piece per plane).  Compile-time tuning constants,
// set from measurements at 8192^2: the fused kernel with 4 columns per lane
// needs more registers than one wave per SIMD allows without spills, with 2 it
// runs three waves per SIMD.  The imaginary-time step (small lattices, a
// wave's row latency bounds it) uses 2: at 2048^2, 20-row segments, 26.4 us
// per step against 29.0 with 4 (256^2: 16.0 against 20.3).
#ifndef TSMI_FP32_VEC
#define TSMI_FP32_VEC 4
#endif
#ifndef TSMI_FP32_VEC2
#define TSMI_FP32_VEC2 2
#endif
#ifndef TSMI_FP32_VEC_IMAG
#define TSMI_FP32_VEC_IMAG 2
#endif
constexpr int kVec32 = TSMI_FP32_VEC;
constexpr int kVec32I = TSMI_FP32_VEC_IMAG;   // the imaginary-time single step
constexpr int kVec32F = TSMI_FP32_VEC2;

}  // namespace


namespace {

size_t plane_bytes(const tsmi_ctx *c) { return (size_t)c->pitch * (size_t)c->Hd * (size_t)c->esz; }

// map_lattice_to_coordinate_space for one axis (src/common.cpp:45-65, Cartesian)
void map_axis(int start, double delta, int n, double length, int count, std::vector<double> &out) {
    out.resize(count);
    const double cc = n * delta * 0.5;
    for (int i = 0; i < count; ++i) {
        double idx = start * delta + 0.5 * delta + i * delta;
        if (idx - cc < -length * 0.5) idx += length;
        if (idx - cc > length * 0.5) idx -= length;
        out[i] = idx - cc;
    }
}

int free_all(tsmi_ctx *c) {
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 2; ++k)
            if (c->psi_base[b][k]) (void)hipFree(c->psi_base[b][k]);
    void *ptrs[] = {c->e_base[0], c->e_base[1], c->e8[0], c->e8[1], c->ph_base, c->ph8, c->flag, c->v_base, c->xc,
                    c->yc, c->div, c->partials, c->part_n, c->obs_partials, c->obs_out, c->ck, c->adj, c->grad};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &k : c->ev)
        for (auto &v : k)
            for (hipEvent_t e : v) (void)hipEventDestroy(e);
    if (c->own) (void)hipStreamDestroy(c->own);
    return TSMI_OK;
}

// Uploads W x H doubles into a pitched plane of the context precision.
int upload_plane(tsmi_ctx *c, void *dst, const double *src) {
    const size_t n = (size_t)c->W * (size_t)c->H;
    if (c->esz == 8) {
        TSMI_HIP(hipMemcpy2DAsync(dst, (size_t)c->pitch * 8, src, (size_t)c->W * 8, (size_t)c->W * 8, c->H,
                                  hipMemcpyHostToDevice, c->stream));
        TSMI_HIP(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];   // rounded once
        TSMI_HIP(hipMemcpy2DAsync(dst, (size_t)c->pitch * 4, tmp.data(), (size_t)c->W * 4, (size_t)c->W * 4, c->H,
                                  hipMemcpyHostToDevice, c->stream));
        TSMI_HIP(hipStreamSynchronize(c->stream));
    }
    return TSMI_OK;
}

// The divisor of a pending imaginary-time renormalisation, computed into c->div
// on the device from the last step's partial sums (nullptr: none pending).
const double *pending_div(tsmi_ctx *c) {
    if (!c->d.imag_time || !c->div_pending) return nullptr;
    hipLaunchKernelGGL(divisor_kernel<>, dim3(1), dim3(kWave), 0, c->stream,
                       (const double *)(c->partials + (size_t)c->div_slot * c->n_waves),
                       (const int *)(c->part_n + c->div_slot), c->d.delta_x * c->d.delta_y, c->d.norm_target, c->div);
    return c->div;
}

int current_div(tsmi_ctx *c, double *out) {
    *out = 1.0;
    if (const double *d = pending_div(c)) {
        TSMI_HIP(hipGetLastError());
        TSMI_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TSMI_HIP(hipStreamSynchronize(c->stream));
    }
    return TSMI_OK;
}

// W x H FP64 (row stride W) -> pitched plane of T, rounding once.
template <typename T>
__global__ void import_plane(const double *__restrict__ src, T *__restrict__ dst, int W, int H, int pitch) {
    const long n = (long)W * H;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long y = i / W, x = i - y * W;
        dst[y * pitch + x] = (T)src[i];
    }
}

// pitched plane of T -> W x H FP64, divided by *div when given
template <typename T>
__global__ void export_plane(const T *__restrict__ src, double *__restrict__ dst, int W, int H, int pitch,
                             const double *__restrict__ div) {
    const long n = (long)W * H;
    const double d = div ? *div : 1.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long y = i / W, x = i - y * W;
        double v = (double)src[y * pitch + x];
        if (div) v /= d;
        dst[i] = v;
    }
}

// E table of the step: E*corr computed in FP64, each component then rounded
// once to the kernel precision.  corr is the step's product of the scaled-frame
// factors (see tsmi_create).  sim may be null (imaginary time: E is real).
// The same pass writes the one-plane form: phi = -arg(E corr) when |E| = 1
// (real time), lambda = -ln(E corr) when E > 0 (imaginary time); a cell without
// that form sets *bad.
template <typename T>
__global__ void import_exp(const double *__restrict__ sre, const double *__restrict__ sim, double cr, double ci,
                           T *__restrict__ dre, T *__restrict__ dim, T *__restrict__ dph, unsigned *bad, int W,
                           int H, int pitch) {
    const long n = (long)W * H;
    bool ok = true;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long y = i / W, x = i - y * W;
        const double er = sre[i], ei = sim ? sim[i] : 0.0;
        const double zr = er * cr - ei * ci, zi = er * ci + ei * cr;
        dre[y * pitch + x] = (T)zr;
        dim[y * pitch + x] = (T)zi;
        double ph = 0.0;
        if (sim) {
            ok = ok && fabs(er * er + ei * ei - 1.0) <= 1e-12;
            ph = -atan2(zi, zr);
        } else {
            ok = ok && zr > 0.0;
            ph = zr > 0.0 ? -log(zr) : 0.0;
        }
        dph[y * pitch + x] = (T)ph;
    }
    if (!ok) atomicOr(bad, 1u);
}

// Imports E (FP64 W x H on the device) and decides whether the steps read its
// one-plane form; syncs to read the import's flag.
int launch_import_exp(tsmi_ctx *c, const double *re, const double *im) {
    TSMI_HIP(hipMemsetAsync(c->flag, 0, sizeof(unsigned), c->stream));
    if (c->esz == 4)
        hipLaunchKernelGGL(import_exp<float>, dim3(2048), dim3(256), 0, c->stream, re, im, c->corr_re, c->corr_im,
                           (float *)c->e_re, (float *)c->e_im, (float *)c->ph, c->flag, c->W, c->H, c->pitch);
    else
        hipLaunchKernelGGL(import_exp<double>, dim3(2048), dim3(256), 0, c->stream, re, im, c->corr_re, c->corr_im,
                           (double *)c->e_re, (double *)c->e_im, (double *)c->ph, c->flag, c->W, c->H, c->pitch);
    TSMI_HIP(hipGetLastError());
    unsigned bad = 1;
    TSMI_HIP(hipMemcpyAsync(&bad, c->flag, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
    TSMI_HIP(hipStreamSynchronize(c->stream));
    c->ph_mode = bad == 0;
    return TSMI_OK;
}

// E = exp(-i dt V) (real time) or exp(-dt V) (imaginary time) of a raw FP64
// potential W x H (row stride W), as Solver::initialize_exp_potential tabulates
// it (src/solver.cpp:74-109), times the frame factor, rounded once; the raw V
// goes to the pitched FP64 plane used by the energies.
template <typename T>
__global__ void exp_from_potential(const double *__restrict__ v, double dt, int imag, double cr, double ci,
                                   T *__restrict__ dre, T *__restrict__ dim, T *__restrict__ dph,
                                   double *__restrict__ vout, int W, int H, int pitch) {
    const long n = (long)W * H;
    // one-plane form: phi = dt V - arg(corr), lambda = dt V - ln(corr)
    const double shift = imag ? log(cr) : atan2(ci, cr);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long y = i / W, x = i - y * W;
        const double vi = v[i];
        double er, ei;
        if (imag) {
            er = exp(-dt * vi);
            ei = 0.0;
        } else {
            sincos(-dt * vi, &ei, &er);
        }
        dre[y * pitch + x] = (T)(er * cr - ei * ci);
        dim[y * pitch + x] = (T)(er * ci + ei * cr);
        dph[y * pitch + x] = (T)(dt * vi - shift);
        vout[y * pitch + x] = vi;
    }
}

int import_device(tsmi_ctx *c, void *dst, const double *src) {
    const int blocks = 2048;
    if (c->esz == 8)
        hipLaunchKernelGGL(import_plane<double>, dim3(blocks), dim3(256), 0, c->stream, src, (double *)dst, c->W, c->H, c->pitch);
    else
        hipLaunchKernelGGL(import_plane<float>, dim3(blocks), dim3(256), 0, c->stream, src, (float *)dst, c->W, c->H, c->pitch);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

// Event pair around a step-kernel launch when timing is on (kind 1: fused).
int time_mark(tsmi_ctx *c, int kind, int which) {
    if (!c->timing) return TSMI_OK;
    std::vector<hipEvent_t> &v = c->ev[kind][which];
    const size_t i = c->ev_used[kind];
    if (i >= v.size()) {
        hipEvent_t e = nullptr;
        TSMI_HIP(hipEventCreate(&e));
        v.push_back(e);
    }
    TSMI_HIP(hipEventRecord(v[i], c->stream));
    if (which == 1) ++c->ev_used[kind];
    return TSMI_OK;
}

// The periodic halo refresh of a state (or E) allocation: the reference's
// 4-wide exchange on the reference tile, or on the 8-wide layout the full wrap
// of all 8 halo cells, whose inner 4 are exactly what the 4-wide exchange writes.
template <typename T>
int exchange_halo(tsmi_ctx *c, void *base_re, void *base_im) {
    if (!c->d.periodic_x && !c->d.periodic_y) return TSMI_OK;
    const int W = c->halo8 ? c->Wd : c->W, H = c->halo8 ? c->Hd : c->H;
    const int hx = c->halo8 ? 2 * c->d.halo_x : c->d.halo_x, hy = c->halo8 ? 2 * c->d.halo_y : c->d.halo_y;
    constexpr int V = 16 / sizeof(T);
    const uintptr_t al = (uintptr_t)base_re | (uintptr_t)base_im;
    if (c->d.periodic_x && c->d.periodic_y && hx % V == 0 && W % V == 0 && c->pitch % V == 0 && al % 16 == 0 &&
        W >= 3 * hx && H >= 3 * hy) {
        // 16-byte pieces (8192^2, 8-wide halo, per exchange: FP32 8.4 -> 7.3 us,
        // FP64 8.0 -> 6.9 us; a launch of this size is bound by its latency)
        const long pieces = exchange_pieces(W, H, hx, hy, V);
        const int eb = (int)std::min<long>(1024, (pieces + 255) / 256);
        hipLaunchKernelGGL((halo_exchange_vec_kernel<T>), dim3(eb), dim3(256), 0, c->stream, (T *)base_re,
                           (T *)base_im, W, H, c->pitch, hx, hy);
        TSMI_HIP(hipGetLastError());
        return TSMI_OK;
    }
    const long cells = 2L * hy * W + 2L * hx * H;
    const int eb = (int)std::min<long>(1024, (cells + 255) / 256);
    hipLaunchKernelGGL((halo_exchange_kernel<T>), dim3(eb), dim3(256), 0, c->stream, (T *)base_re, (T *)base_im, W, H,
                       c->pitch, c->d.periodic_x, c->d.periodic_y, hx, hy);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

// Segments of the fused pass sized to one round of resident waves of the
// instance that runs (its registers, hence its occupancy, depend on the scaled
// form and on the E form); recomputed when either changes.
// Resident waves of a kernel on the current device at its occupancy (0 when
// there is no device to ask) and the device's SIMD count.
template <typename K>
long resident_waves(K kernel, long *simds) {
    int dev = 0, cus = 0, per_cu = 0;
    *simds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || cus <= 0 ||
        per_cu <= 0) {
        (void)hipGetLastError();
        return 0;
    }
    *simds = 4L * cus;
    return (long)cus * per_cu * (kBlock / kWave);
}

// Geometry of the two-step pass over the allocation (Wd x Hd): strips of
// kWave*VEC - 16 output columns, segments of one round of resident waves, at
// least 16 rows -- 12 where 16-row segments leave SIMDs without a wave: a
// wave's row latency then bounds the pass (measured, 1024^2 closed, 2000 steps:
// FP32 33.1 -> 30.3 ms, FP64 38.1 -> 34.9 ms).
void plan_fused(int Wd, int Hd, long long pitch, int tile_rows, int vec, long resident, long simds, int *strips,
                int *rows, int *segs) {
    const int sw2 = kWave * vec - 4 * kApron;
    *strips = (Wd + sw2 - 1) / sw2;
    *rows = plan_seg_rows(Hd, *strips, resident, simds, 16, true, tile_rows, pitch, 4 * kApron);
    *segs = (Hd + *rows - 1) / *rows;
}

template <typename T, int VEC, bool LHY, int MODE, bool PH>
int size_fused(tsmi_ctx *c) {
    const int key = (MODE + 1) * 2 + (PH ? 1 : 0);
    if (c->geo2_key == key) return TSMI_OK;
    long simds = 0;
    const long resident = resident_waves(step2_kernel<T, VEC, LHY, MODE, PH>, &simds);
    plan_fused(c->Wd, c->Hd, c->pitch, c->d.tile_rows, VEC, resident, simds, &c->n_strips2, &c->seg_rows2,
               &c->n_segs2);
    c->n_waves2 = c->n_strips2 * c->n_segs2;
    c->geo2_key = key;
    return TSMI_OK;
}

// Two real-time steps in one pass over the whole allocation (fused_kernel.hpp).
static_assert(kBlock == kStep2Block && kBlock == kStepBlock, "the step kernels' workgroup size is their launch bound");
template <typename T, int VEC, bool LHY, int MODE, bool PH>
int launch_step2(tsmi_ctx *c) {
    if (int st = size_fused<T, VEC, LHY, MODE, PH>(c)) return st;
    StepParams<T> p;
    std::memset(&p, 0, sizeof(p));
    const int nb = 1 - c->cur;
    p.in_re = (const T *)c->psi_base[c->cur][0];
    p.in_im = (const T *)c->psi_base[c->cur][1];
    p.out_re = (T *)c->psi_base[nb][0];
    p.out_im = (T *)c->psi_base[nb][1];
    p.e_re = (const T *)(c->halo8 ? c->e8[0] : c->e_base[0]);
    p.e_im = (const T *)(c->halo8 ? c->e8[1] : c->e_base[1]);
    p.ph = (const T *)(c->halo8 ? c->ph8 : c->ph_base);
    p.mag = c->corr_abs;
    p.W = c->Wd; p.H = c->Hd; p.pitch = c->pitch;
    p.n_strips = c->n_strips2; p.n_waves = c->n_waves2; p.seg_rows = c->seg_rows2;
    p.mH = (T)c->mH; p.tH = (T)c->tH; p.mV = (T)c->mV; p.tV = (T)c->tV;
    p.unH = (T)c->unH; p.unV = (T)c->unV;
    p.c_nl = (T)c->c_nl_s; p.c_lhy = (T)c->c_lhy_s;
    const int blocks = (c->n_waves2 + kBlock / kWave - 1) / (kBlock / kWave);
    if (int st = time_mark(c, 1, 0)) return st;
    if (sizeof(T) == 8 && c->nt_out)
        hipLaunchKernelGGL((step2_kernel<T, VEC, LHY, MODE, PH, sizeof(T) == 8>), dim3(blocks), dim3(kBlock), 0,
                           c->stream, p);
    else
        hipLaunchKernelGGL((step2_kernel<T, VEC, LHY, MODE, PH>), dim3(blocks), dim3(kBlock), 0, c->stream, p);
    TSMI_HIP(hipGetLastError());
    if (int st = time_mark(c, 1, 1)) return st;
    c->cur = nb;
    return TSMI_OK;
}

template <typename T, int VEC, bool LHY, bool PH>
int step2_form(tsmi_ctx *c) {
    if (c->mode == 0) return launch_step2<T, VEC, LHY, 0, PH>(c);
    if (c->mode == 1) return launch_step2<T, VEC, LHY, 1, PH>(c);
    return launch_step2<T, VEC, LHY, -1, PH>(c);
}

template <typename T, int VEC, bool LHY>
int step2_mode(tsmi_ctx *c) {
    return c->ph_mode ? step2_form<T, VEC, LHY, true>(c) : step2_form<T, VEC, LHY, false>(c);
}

int step2_once(tsmi_ctx *c) {
    const bool lhy = c->c_lhy != 0.0;
    if (c->esz == 4) return lhy ? step2_mode<float, kVec32F, true>(c) : step2_mode<float, kVec32F, false>(c);
    return lhy ? step2_mode<double, 2, true>(c) : step2_mode<double, 2, false>(c);
}

template <typename T, bool IMAG, int VEC, bool LHY, int MODE, bool PH>
int launch_step(tsmi_ctx *c, int exchange) {
    StepParams<T> p;
    std::memset(&p, 0, sizeof(p));
    const int nb = 1 - c->cur;
    p.in_re = (const T *)c->psi_re[c->cur];
    p.in_im = (const T *)c->psi_im[c->cur];
    p.out_re = (T *)c->psi_re[nb];
    p.out_im = (T *)c->psi_im[nb];
    p.e_re = (const T *)c->e_re;
    p.e_im = (const T *)c->e_im;
    p.ph = (const T *)c->ph;
    p.mag = IMAG ? 1.0 : c->corr_abs;
    p.W = c->W; p.H = c->H; p.pitch = c->pitch;
    p.n_strips = c->n_strips; p.n_waves = c->n_waves; p.seg_rows = c->seg_rows;
    p.mH = (T)c->mH; p.tH = (T)c->tH; p.mV = (T)c->mV; p.tV = (T)c->tV;
    p.unH = (T)c->unH; p.unV = (T)c->unV;
    p.c_nl = (T)c->c_nl_s; p.c_lhy = (T)c->c_lhy_s;
    p.ix0 = c->ix0; p.ix1 = c->ix1; p.iy0 = c->iy0; p.iy1 = c->iy1;
    if (IMAG) {
        const size_t nw = (size_t)c->n_waves;
        p.part_in = c->div_pending ? c->partials + (size_t)c->div_slot * nw : nullptr;
        p.n_in = c->part_n + c->div_slot;
        p.part_out = c->partials + (size_t)(1 - c->div_slot) * nw;
        p.n_out = c->part_n + (1 - c->div_slot);
        p.norm_target = c->d.norm_target;
        p.dxdy = c->d.delta_x * c->d.delta_y;
    }
    // the grid covers every (strip, segment) pair exactly once
    const int blocks = (c->n_waves + kBlock / kWave - 1) / (kBlock / kWave);
    if (int st = time_mark(c, 0, 0)) return st;
    if (!IMAG && sizeof(T) == 8 && c->nt_out)
        hipLaunchKernelGGL((step_kernel<T, IMAG, VEC, LHY, MODE, PH, !IMAG && sizeof(T) == 8>), dim3(blocks),
                           dim3(kBlock), 0, c->stream, p);
    else
        hipLaunchKernelGGL((step_kernel<T, IMAG, VEC, LHY, MODE, PH>), dim3(blocks), dim3(kBlock), 0, c->stream, p);
    TSMI_HIP(hipGetLastError());
    if (int st = time_mark(c, 0, 1)) return st;
    c->cur = nb;
    if (exchange) {
        const int st = exchange_halo<T>(c, c->psi_base[nb][0], c->psi_base[nb][1]);
        if (st) return st;
    }
    if (IMAG) {
        c->div_slot = 1 - c->div_slot;
        c->div_pending = true;
    }
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY>
hipError_t occupancy_of(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, step_kernel<T, IMAG, VEC, LHY, 0, true>, kBlock, 0);
}

hipError_t occupancy_blocks(int precision, bool imag, bool lhy, int *blocks) {
    if (precision == 32) {
        if (imag) return lhy ? occupancy_of<float, true, kVec32I, true>(blocks) : occupancy_of<float, true, kVec32I, false>(blocks);
        return lhy ? occupancy_of<float, false, kVec32, true>(blocks) : occupancy_of<float, false, kVec32, false>(blocks);
    }
    if (imag) return lhy ? occupancy_of<double, true, 2, true>(blocks) : occupancy_of<double, true, 2, false>(blocks);
    return lhy ? occupancy_of<double, false, 2, true>(blocks) : occupancy_of<double, false, 2, false>(blocks);
}

// Rows per segment of the one-step pass.  One round of resident wavefronts: as
// many row segments per strip as the device holds at this kernel's occupancy,
// so no partial second round idles most of the GPU; on small lattices at least
// 20 rows, below which the 4-row aprons and the per-wave partial sums cost more
// than the extra waves bring -- unless 20-row segments leave SIMDs without a
// wave, where a wave's row latency bounds the step and 12-row segments are
// faster (measured, imaginary time: 1024^2 FP32 28.9 -> 22.5 us, FP64 30.3 ->
// 23.2; at 2048^2 20-row segments against 24: FP32 26.4 vs 28.0, FP64 44.0 vs
// 45.8, and against 12: 27.3 and 52.2).
int plan_single(const tsmi_desc &d, int H, long long pitch, int strips) {
    long resident = 0, simds = 0;
    int dev = 0, cus = 0, blocks_per_cu = 0;
    if (d.tile_rows <= 0 && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        occupancy_blocks(d.precision, d.imag_time != 0, d.coupling_lhy != 0.0, &blocks_per_cu) == hipSuccess &&
        cus > 0 && blocks_per_cu > 0) {
        resident = (long)cus * blocks_per_cu * (kBlock / kWave);
        simds = 4L * cus;
    }
    (void)hipGetLastError();
    return plan_seg_rows(H, strips, resident, simds, 20, true, d.tile_rows, pitch, 2 * kApron);
}

template <typename T, bool IMAG, int VEC, bool LHY, bool PH>
int step_form(tsmi_ctx *c, int exchange) {
    if constexpr (IMAG) {
        return launch_step<T, IMAG, VEC, LHY, 0, PH>(c, exchange);   // form A on both axes
    } else {
        if (c->mode == 0) return launch_step<T, IMAG, VEC, LHY, 0, PH>(c, exchange);
        if (c->mode == 1) return launch_step<T, IMAG, VEC, LHY, 1, PH>(c, exchange);
        return launch_step<T, IMAG, VEC, LHY, -1, PH>(c, exchange);
    }
}

template <typename T, bool IMAG, int VEC, bool LHY>
int step_mode(tsmi_ctx *c, int exchange) {
    return c->ph_mode ? step_form<T, IMAG, VEC, LHY, true>(c, exchange)
                      : step_form<T, IMAG, VEC, LHY, false>(c, exchange);
}

template <typename T, int VEC, int VEC_IMAG>
int step_typed(tsmi_ctx *c, int exchange) {
    // the LHY term is compiled in only when g_LHY != 0 (it is exactly zero otherwise)
    const bool lhy = c->c_lhy != 0.0;
    if (c->d.imag_time)
        return lhy ? step_mode<T, true, VEC_IMAG, true>(c, exchange) : step_mode<T, true, VEC_IMAG, false>(c, exchange);
    return lhy ? step_mode<T, false, VEC, true>(c, exchange) : step_mode<T, false, VEC, false>(c, exchange);
}

int step_once(tsmi_ctx *c, int exchange) {
    return c->esz == 4 ? step_typed<float, kVec32, kVec32I>(c, exchange) : step_typed<double, 2, 2>(c, exchange);
}

int exchange_planes(tsmi_ctx *c, void *re, void *im) {
    return c->esz == 4 ? exchange_halo<float>(c, re, im) : exchange_halo<double>(c, re, im);
}

// Solver::evolve(n) for one component (src/solver.cpp:186-230): n steps, the
// halo exchange after all but the last (after the last too when exchange_last).
// With two steps per pass, on closed axes the pairs simply replace single
// steps; on periodic axes (8-wide halo layout) the first step runs alone on the
// reference tile (its halo is whatever the caller or the previous call left, as
// in the reference), every pass starts from a freshly wrapped halo, and the last
// step runs alone so that the halo it leaves is the reference's unrefreshed one
// (or the refreshed one, when exchange_last).
int evolve_steps(tsmi_ctx *c, int n, bool exchange_last = false) {
    int st = TSMI_OK;
    const int last_ex = exchange_last ? 1 : 0;
    if (c->sweeps != 2 || n < 2 || (c->halo8 && n < 3)) {
        for (int i = 0; i < n && !st; ++i) st = step_once(c, i != n - 1 ? 1 : last_ex);
        return st;
    }
    if (!c->halo8) {
        for (int i = 0; i + 1 < n && !st; i += 2) st = step2_once(c);
        if (!st && (n & 1)) st = step_once(c, last_ex);
        return st;
    }
    st = step_once(c, 1);
    int m = n - 2;
    for (; m >= 2 && !st; m -= 2) {
        st = step2_once(c);
        if (!st) st = exchange_planes(c, c->psi_base[c->cur][0], c->psi_base[c->cur][1]);
    }
    if (!st && m == 1) st = step_once(c, 1);
    if (!st) st = step_once(c, last_ex);
    return st;
}

// The fused passes read E over the whole 8-wide halo: the reference-tile table
// with every halo cell replaced by its wrapped inner cell.
int wrap_exp(tsmi_ctx *c) {
    if (!c->halo8) return TSMI_OK;
    const size_t pb = plane_bytes(c);
    TSMI_HIP(hipMemcpyAsync(c->e8[0], c->e_base[0], pb, hipMemcpyDeviceToDevice, c->stream));
    TSMI_HIP(hipMemcpyAsync(c->e8[1], c->e_base[1], pb, hipMemcpyDeviceToDevice, c->stream));
    TSMI_HIP(hipMemcpyAsync(c->ph8, c->ph_base, pb, hipMemcpyDeviceToDevice, c->stream));
    if (int st = exchange_planes(c, c->e8[0], c->e8[1])) return st;
    return exchange_planes(c, c->ph8, c->ph8);   // one plane: the copy runs twice, to the same values
}

// One reverse step: (s_in, g_in) -> (s_out, g_out), gradient into c->grad.
template <typename T, int VEC, bool NL>
int launch_adjoint(tsmi_ctx *c, const void *s_re, const void *s_im, const void *g_re, const void *g_im,
                   void *so_re, void *so_im, void *go_re, void *go_im) {
    AdjParams<T> p;
    std::memset(&p, 0, sizeof(p));
    p.s_in_re = (const T *)s_re; p.s_in_im = (const T *)s_im;
    p.g_in_re = (const T *)g_re; p.g_in_im = (const T *)g_im;
    p.s_out_re = (T *)so_re; p.s_out_im = (T *)so_im;
    p.g_out_re = (T *)go_re; p.g_out_im = (T *)go_im;
    p.e_re = (const T *)c->e_re; p.e_im = (const T *)c->e_im;
    p.grad = c->grad;
    p.W = c->W; p.H = c->H; p.pitch = c->pitch;
    p.n_strips = c->n_strips; p.n_waves = c->n_waves; p.seg_rows = c->seg_rows;
    p.mH = (T)c->amH; p.tH = (T)c->atH; p.mV = (T)c->amV; p.tV = (T)c->atV;
    p.unH = (T)c->aunH; p.unV = (T)c->aunV;
    p.c_nl = (T)c->c_nl_s; p.c_lhy = (T)c->c_lhy_s;
    p.gscale = c->d.delta_t * c->corr_abs;
    const int blocks = (c->n_waves + kBlock / kWave - 1) / (kBlock / kWave);
    hipLaunchKernelGGL((adjoint_kernel<T, VEC, NL>), dim3(blocks), dim3(kBlock), 0, c->stream, p);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

int adjoint_step(tsmi_ctx *c, const void *s_re, const void *s_im, const void *g_re, const void *g_im,
                 void *so_re, void *so_im, void *go_re, void *go_im) {
    const bool nl = c->c_nl != 0.0 || c->c_lhy != 0.0;
    if (c->esz == 4)
        return nl ? launch_adjoint<float, 4, true>(c, s_re, s_im, g_re, g_im, so_re, so_im, go_re, go_im)
                  : launch_adjoint<float, 4, false>(c, s_re, s_im, g_re, g_im, so_re, so_im, go_re, go_im);
    return nl ? launch_adjoint<double, 2, true>(c, s_re, s_im, g_re, g_im, so_re, so_im, go_re, go_im)
              : launch_adjoint<double, 2, false>(c, s_re, s_im, g_re, g_im, so_re, so_im, go_re, go_im);
}

// The halo exchange, its transpose, and the clearing of the halo, on a W x H
// field of the reference tile's layout (origin at `re`, row pitch c->pitch):
// the reverse sweep's view of a periodic lattice, whatever the allocation.
template <typename T>
int view_exchange(tsmi_ctx *c, void *re, void *im) {
    const long cells = exchange_cells(c->W, c->H, c->d.periodic_x, c->d.periodic_y, c->d.halo_x, c->d.halo_y);
    if (cells == 0) return TSMI_OK;
    const int eb = (int)std::min<long>(1024, (cells + 255) / 256);
    hipLaunchKernelGGL((halo_exchange_kernel<T>), dim3(eb), dim3(256), 0, c->stream, (T *)re, (T *)im, c->W, c->H,
                       c->pitch, c->d.periodic_x, c->d.periodic_y, c->d.halo_x, c->d.halo_y);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

template <typename T>
int view_fold(tsmi_ctx *c, void *re, void *im) {
    for (int stage = 0; stage < 2; ++stage) {
        const long cells = fold_cells(c->W, c->H, c->d.periodic_x, c->d.periodic_y, c->d.halo_x, c->d.halo_y, stage);
        if (cells == 0) continue;
        const int eb = (int)std::min<long>(1024, (cells + 255) / 256);
        hipLaunchKernelGGL((halo_fold_kernel<T>), dim3(eb), dim3(256), 0, c->stream, (T *)re, (T *)im, c->W, c->H,
                           c->pitch, c->d.periodic_x, c->d.periodic_y, c->d.halo_x, c->d.halo_y, stage);
        TSMI_HIP(hipGetLastError());
    }
    return TSMI_OK;
}

template <typename T>
int view_zero_halo(tsmi_ctx *c, void *p) {
    const long n = (long)c->W * c->H;
    const int eb = (int)std::min<long>(2048, (n + 255) / 256);
    hipLaunchKernelGGL((zero_outside_kernel<T>), dim3(eb), dim3(256), 0, c->stream, (T *)p, c->W, c->H, c->pitch,
                       c->ix0, c->ix1, c->iy0, c->iy1);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

int view_exchange_any(tsmi_ctx *c, void *re, void *im) {
    return c->esz == 4 ? view_exchange<float>(c, re, im) : view_exchange<double>(c, re, im);
}

// The reverse sweep over the steps of the last checkpointed evolve.  g_re/g_im
// are FP64 W x H planes on the device; the gradient is left in c->grad and the
// adjoint of the initial state in the returned planes.
//
// Periodic axes: the tile with a refreshed halo evolves as the periodic lattice
// (the halo is as wide as one step's reach), so the sweep runs on the periodic
// lattice: dL/dpsi_N's halo entries are folded onto the cells they mirror and
// the halo then holds their copies, psi_N gets its halo refreshed (the forward's
// last step leaves it stale), and after every reverse step both states get
// their halo refreshed.  The gradient and dL/dpsi_0 are those of the inner
// cells -- each with respect to the value a cell shares with its halo copies --
// and zero on the halo.
int reverse_sweep(tsmi_ctx *c, const double *g_re, const double *g_im, void **g0_re, void **g0_im) {
    const size_t pb = plane_bytes(c);
    const size_t pd = (size_t)c->pitch * (size_t)c->Hd * sizeof(double);
    const size_t view_off = c->off_el * (size_t)c->esz;
    const bool per = c->d.periodic_x || c->d.periodic_y;
    if (!c->adj) TSMI_HIP(hipMalloc(&c->adj, 8 * pb));
    if (!c->grad) TSMI_HIP(hipMalloc((void **)&c->grad, pd));
    char *a = (char *)c->adj;
    void *L[2][2] = {{a, a + pb}, {a + 2 * pb, a + 3 * pb}};
    void *S[2][2] = {{a + 4 * pb, a + 5 * pb}, {a + 6 * pb, a + 7 * pb}};
    // rows are padded; the padding of every plane must read as zero
    TSMI_HIP(hipMemsetAsync(c->adj, 0, 8 * pb, c->stream));
    TSMI_HIP(hipMemsetAsync(c->grad, 0, pd, c->stream));
    int st = import_device(c, L[0][0], g_re);
    if (!st) st = import_device(c, L[0][1], g_im);
    if (st) return st;
    const void *s_re = c->psi_re[c->cur], *s_im = c->psi_im[c->cur];
    if (per) {
        st = c->esz == 4 ? view_fold<float>(c, L[0][0], L[0][1]) : view_fold<double>(c, L[0][0], L[0][1]);
        if (!st) st = view_exchange_any(c, L[0][0], L[0][1]);
        const size_t row = (size_t)c->pitch * c->esz;
        TSMI_HIP(hipMemcpy2DAsync(S[0][0], row, s_re, row, (size_t)c->W * c->esz, c->H, hipMemcpyDeviceToDevice,
                                  c->stream));
        TSMI_HIP(hipMemcpy2DAsync(S[0][1], row, s_im, row, (size_t)c->W * c->esz, c->H, hipMemcpyDeviceToDevice,
                                  c->stream));
        if (!st) st = view_exchange_any(c, S[0][0], S[0][1]);
        if (st) return st;
        s_re = S[0][0];
        s_im = S[0][1];
    }
    int k = 0;
    for (int n = c->ck_steps - 1; n >= 0; --n) {
        const int o = 1 - k;
        st = adjoint_step(c, s_re, s_im, L[k][0], L[k][1], S[o][0], S[o][1], L[o][0], L[o][1]);
        if (!st && per) st = view_exchange_any(c, L[o][0], L[o][1]);
        if (st) return st;
        if (n % c->ck_every == 0) {   // psi_n is a checkpoint: use it instead of the reversed state
            const char *ck = (const char *)c->ck + (size_t)(n / c->ck_every) * 2 * pb + view_off;
            s_re = ck;
            s_im = ck + pb;
        } else {
            if (per && (st = view_exchange_any(c, S[o][0], S[o][1]))) return st;
            s_re = S[o][0];
            s_im = S[o][1];
        }
        k = o;
    }
    if (per) {
        st = view_zero_halo<double>(c, c->grad);
        if (!st) st = c->esz == 4 ? view_zero_halo<float>(c, L[k][0]) : view_zero_halo<double>(c, L[k][0]);
        if (!st) st = c->esz == 4 ? view_zero_halo<float>(c, L[k][1]) : view_zero_halo<double>(c, L[k][1]);
        if (st) return st;
    }
    *g0_re = L[k][0];
    *g0_im = L[k][1];
    return TSMI_OK;
}

// pitched plane of the context precision -> W x H FP64 on the device
int export_device(tsmi_ctx *c, const void *src, double *dst) {
    const int blocks = 2048;
    if (c->esz == 8)
        hipLaunchKernelGGL(export_plane<double>, dim3(blocks), dim3(256), 0, c->stream, (const double *)src, dst,
                           c->W, c->H, c->pitch, (const double *)nullptr);
    else
        hipLaunchKernelGGL(export_plane<float>, dim3(blocks), dim3(256), 0, c->stream, (const float *)src, dst,
                           c->W, c->H, c->pitch, (const double *)nullptr);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

// Geometry of an observables pass over a tile of W x H cells (row pitch in
// elements): the inner region and the 2-cell stencil margin next to closed
// edges (src/solver.cpp:356-359, src/model.cpp:424-427).
struct ObsGeom {
    int W, H, pitch, x0, x1, y0, y1, kx0, kx1, ky0, ky1;
};

ObsGeom obs_geom(const tsmi_desc &d, int W, int H, int pitch) {
    ObsGeom g;
    g.W = W; g.H = H; g.pitch = pitch;
    g.x0 = d.periodic_x ? d.halo_x : 0; g.x1 = W - g.x0;
    g.y0 = d.periodic_y ? d.halo_y : 0; g.y1 = H - g.y0;
    g.kx0 = g.x0 + (d.periodic_x ? 0 : 2); g.kx1 = g.x1 - (d.periodic_x ? 0 : 2);
    g.ky0 = g.y0 + (d.periodic_y ? 0 : 2); g.ky1 = g.y1 - (d.periodic_y ? 0 : 2);
    return g;
}

// Grid of a pass: one block per kObsBlockCols columns of the inner region and row bands
// of at least 16 rows, at most kObsBlocksMax blocks.  Returns false when the
// inner region is wider than kObsBlocksMax * kObsBlockCols columns.
bool obs_grid(const ObsGeom &g, dim3 *grid, int *band) {
    const int cb = (g.x1 - g.x0 + kObsBlockCols - 1) / kObsBlockCols;
    if (cb < 1 || cb > kObsBlocksMax) return false;
    const int rows = std::max(1, g.y1 - g.y0);
    const int rb_max = kObsBlocksMax / cb;
    *band = std::max(16, (rows + rb_max - 1) / rb_max);
    *grid = dim3(cb, (rows + *band - 1) / *band);
    return true;
}

// One pass of kind `kind` (ObsKind) over the planes, the block partials in
// `partials` (kObsBlocksMax * OBS_COUNT doubles), then obs_finalize_kernel,
// which writes kObsOut doubles to `out` (device).  `div` (device, may be null)
// holds a pending imaginary-time divisor of the stored state.  No allocation,
// no synchronisation.
template <typename T>
int launch_observables(const tsmi_desc &d, const ObsGeom &g, const T *re, const T *im, const double *v,
                       const double *xc, const double *yc, int kind, bool rows_free, double *partials,
                       const double *div, double *out, hipStream_t stream, const char *who) {
    dim3 grid;
    int band = 0;
    if (!obs_grid(g, &grid, &band))
        return fail(TSMI_E_INVALID, std::string(who) + ": the inner region is wider than " +
                                        std::to_string(kObsBlocksMax * kObsBlockCols) + " columns");
    ObsParams<T> p;
    std::memset(&p, 0, sizeof(p));
    p.re = re; p.im = im; p.v = v; p.xc = xc; p.yc = yc;
    p.W = g.W; p.H = g.H; p.pitch = g.pitch;
    p.x0 = g.x0; p.x1 = g.x1; p.y0 = g.y0; p.y1 = g.y1;
    p.kx0 = g.kx0; p.kx1 = g.kx1; p.ky0 = g.ky0; p.ky1 = g.ky1;
    p.rows_free = rows_free ? 1 : 0;
    p.band = band;
    p.dx = d.delta_x; p.dy = d.delta_y;
    p.partials = partials;
    const bool lhy = kind != OBS_KIND_NORM && d.coupling_lhy != 0.0;
    if (kind == OBS_KIND_NORM)
        hipLaunchKernelGGL((observables_kernel<T, OBS_KIND_NORM, false>), grid, dim3(kObsThreads), 0, stream, p);
    else if (kind == OBS_KIND_ENERGY && lhy)
        hipLaunchKernelGGL((observables_kernel<T, OBS_KIND_ENERGY, true>), grid, dim3(kObsThreads), 0, stream, p);
    else if (kind == OBS_KIND_ENERGY)
        hipLaunchKernelGGL((observables_kernel<T, OBS_KIND_ENERGY, false>), grid, dim3(kObsThreads), 0, stream, p);
    else if (lhy)
        hipLaunchKernelGGL((observables_kernel<T, OBS_KIND_ALL, true>), grid, dim3(kObsThreads), 0, stream, p);
    else
        hipLaunchKernelGGL((observables_kernel<T, OBS_KIND_ALL, false>), grid, dim3(kObsThreads), 0, stream, p);
    TSMI_HIP(hipGetLastError());
    ObsFinal f;
    std::memset(&f, 0, sizeof(f));
    f.kind = kind;
    f.nq = kind == OBS_KIND_NORM ? 1 : kind == OBS_KIND_ENERGY ? OBS_LHY + 1 : OBS_COUNT;
    f.blocks = (int)(grid.x * grid.y);
    f.rows2d = g.H > 1;
    f.div = div;
    f.mass = d.mass; f.g = d.coupling; f.g_lhy = d.coupling_lhy;
    f.dx = d.delta_x; f.dy = d.delta_y; f.length_x = d.length_x; f.dim_x = d.dim_x;
    hipLaunchKernelGGL(obs_finalize_kernel, dim3(1), dim3(64 * OBS_COUNT), 0, stream, (const double *)partials, f,
                       out);
    TSMI_HIP(hipGetLastError());
    return TSMI_OK;
}

// The pass over the context's current state; out: kObsOut doubles (host).
int observables(tsmi_ctx *c, int kind, bool rows_free, double out[kObsOut]) {
    const ObsGeom g = obs_geom(c->d, c->W, c->H, c->pitch);
    const double *div = pending_div(c);
    const double *v = c->has_v ? c->v : nullptr;
    const int st = c->esz == 4
        ? launch_observables<float>(c->d, g, (const float *)c->psi_re[c->cur], (const float *)c->psi_im[c->cur], v,
                                    c->xc, c->yc, kind, rows_free, c->obs_partials, div, c->obs_out, c->stream,
                                    "observables")
        : launch_observables<double>(c->d, g, (const double *)c->psi_re[c->cur], (const double *)c->psi_im[c->cur],
                                     v, c->xc, c->yc, kind, rows_free, c->obs_partials, div, c->obs_out, c->stream,
                                     "observables");
    if (st) return st;
    TSMI_HIP(hipMemcpyAsync(out, c->obs_out, kObsOut * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TSMI_HIP(hipStreamSynchronize(c->stream));
    return TSMI_OK;
}

}  // namespace

int tsmi::ctx_exchange(tsmi_ctx *c, void *re, void *im) { return exchange_planes(c, re, im); }

extern "C" {

const char *tsmi_last_error(void) { return g_err.c_str(); }

int tsmi_create(const tsmi_desc *desc, tsmi_ctx **out) {
    if (!desc || !out) return fail(TSMI_E_INVALID, "tsmi_create: null argument");
    *out = nullptr;
    const tsmi_desc &d = *desc;
    if (int st = check_desc(d, "tsmi_create")) return st;
    const long long W = (long long)d.dim_x + (d.periodic_x ? 2 * d.halo_x : 0);
    const long long H = (long long)d.dim_y + (d.periodic_y ? 2 * d.halo_y : 0);
    if (d.sweeps < 0 || d.sweeps > 2) return fail(TSMI_E_INVALID, "tsmi_create: sweeps must be 0, 1 or 2");
    // two steps per pass: real time, both axes closed or both periodic
    const bool fusable = !d.imag_time && ((d.periodic_x && d.periodic_y) || (!d.periodic_x && !d.periodic_y));
    if (d.sweeps == 2 && !fusable)
        return fail(TSMI_E_UNSUPPORTED,
                    "tsmi_create: two steps per pass need real time and both axes closed or both periodic");
    // 0 chooses two steps per pass wherever they are exact
    const int sweeps = (d.sweeps != 1 && fusable) ? 2 : 1;
    const bool halo8 = sweeps == 2 && d.periodic_x && d.periodic_y;
    const long long Wd = W + (halo8 ? 2 * d.halo_x : 0), Hd = H + (halo8 ? 2 * d.halo_y : 0);
    const long long pitch = (Wd + 63) / 64 * 64;

    tsmi_ctx *c = new tsmi_ctx();
    c->d = d;
    c->W = (int)W; c->H = (int)H; c->pitch = (int)pitch;
    c->Wd = (int)Wd; c->Hd = (int)Hd;
    c->halo8 = halo8;
    c->off_el = halo8 ? (size_t)d.halo_y * (size_t)pitch + (size_t)d.halo_x : 0;
    // FP64 real time stores its output rows non-temporal when a step's rows
    // (state in and out, phase plane: five planes) exceed the 256 MB Infinity
    // Cache of the MI355X, which then cannot return them to the next step
    {
        const long long plane = pitch * Hd * (d.precision == 32 ? 4 : 8);
        c->nt_out = d.precision == 64 && !d.imag_time && 5 * plane > (256LL << 20);
    }
    c->sweeps = sweeps;
    c->esz = d.precision == 32 ? 4 : 8;
    c->vec = c->esz == 4 ? (d.imag_time ? kVec32I : kVec32) : 2;
    const int sw = kWave * c->vec - 2 * kApron;
    c->n_strips = (c->W + sw - 1) / sw;
    c->seg_rows = plan_single(d, c->H, pitch, c->n_strips);
    c->n_segs = (c->H + c->seg_rows - 1) / c->seg_rows;
    c->n_waves = c->n_strips * c->n_segs;
    c->ix0 = d.periodic_x ? d.halo_x : 0; c->ix1 = c->W - c->ix0;
    c->iy0 = d.periodic_y ? d.halo_y : 0; c->iy1 = c->H - c->iy0;
    // kinetic and nonlinear coefficients in FP64 (src/cpukernel.cpp:195-216) and
    // the kernel's scaled-frame factors
    const StepCoefs k = step_coefs(d, c->H, c->esz);
    c->aH = k.aH; c->bH = k.bH; c->aV = k.aV; c->bV = k.bV;
    c->c_nl = k.c_nl; c->c_lhy = k.c_lhy;
    c->mH = k.mH; c->tH = k.tH; c->unH = k.unH;
    c->mV = k.mV; c->tV = k.tV; c->unV = k.unV;
    c->mode = k.mode;
    c->corr_re = k.corr_re; c->corr_im = k.corr_im;
    c->c_nl_s = k.c_nl_s; c->c_lhy_s = k.c_lhy_s;
    // the reverse sweep: real time; on a periodic axis an inner extent of at
    // least twice the halo, and no x-periodic tile whose exchange leaves rows
    // unrefreshed (closed y with a y halo: the reference refreshes H - 2 halo_y
    // rows), which would not evolve as a periodic lattice
    c->adj_ok = !d.imag_time && (!d.periodic_x || d.dim_x >= 2 * d.halo_x) &&
                (!d.periodic_y || d.dim_y >= 2 * d.halo_y) && !(d.periodic_x && !d.periodic_y && d.halo_y > 0);
    c->amH = k.amH; c->atH = k.atH; c->aunH = k.aunH;
    c->amV = k.amV; c->atV = k.atV; c->aunV = k.aunV;
    c->corr_abs = k.corr_abs;

    auto bail = [&](int code, const std::string &m) {
        free_all(c);
        delete c;
        return fail(code, m);
    };
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(TSMI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    c->stream = c->own;
    const size_t pb = plane_bytes(c);
    const size_t pd = (size_t)c->pitch * (size_t)c->Hd * sizeof(double);
    c->obs_blocks = std::min(kObsBlocksMax, c->H);
    struct { void **p; size_t n; } allocs[] = {
        {&c->psi_base[0][0], pb}, {&c->psi_base[0][1], pb}, {&c->psi_base[1][0], pb}, {&c->psi_base[1][1], pb},
        {&c->e_base[0], pb}, {&c->e_base[1], pb}, {&c->e8[0], halo8 ? pb : 0}, {&c->e8[1], halo8 ? pb : 0},
        {&c->ph_base, pb}, {&c->ph8, halo8 ? pb : 0}, {(void **)&c->flag, 64},
        {(void **)&c->v_base, pd},
        {(void **)&c->xc, (size_t)c->W * sizeof(double)}, {(void **)&c->yc, (size_t)c->H * sizeof(double)},
        {(void **)&c->div, 2 * sizeof(double)}, {(void **)&c->partials, 2 * (size_t)c->n_waves * sizeof(double)},
        {(void **)&c->part_n, 64}, {(void **)&c->obs_partials, (size_t)kObsBlocksMax * OBS_COUNT * sizeof(double)},
        {(void **)&c->obs_out, kObsOut * sizeof(double)},
    };
    for (auto &a : allocs) {
        if (a.n == 0) continue;
        e = hipMalloc(a.p, a.n);
        if (e != hipSuccess) return bail(TSMI_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemsetAsync(*a.p, 0, a.n, c->stream);
        if (e != hipSuccess) return bail(TSMI_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    const size_t ob = c->off_el * (size_t)c->esz;
    for (int b = 0; b < 2; ++b) {
        c->psi_re[b] = (char *)c->psi_base[b][0] + ob;
        c->psi_im[b] = (char *)c->psi_base[b][1] + ob;
    }
    c->e_re = (char *)c->e_base[0] + ob;
    c->e_im = (char *)c->e_base[1] + ob;
    c->ph = (char *)c->ph_base + ob;
    c->v = c->v_base + c->off_el;
    std::vector<double> xs, ys;
    map_axis(d.periodic_x ? -d.halo_x : 0, d.delta_x, d.dim_x, d.length_x, c->W, xs);
    map_axis(d.periodic_y ? -d.halo_y : 0, d.delta_y, d.dim_y, d.length_y, c->H, ys);
    e = hipMemcpyAsync(c->xc, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->yc, ys.data(), ys.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(TSMI_E_HIP, std::string("coordinate upload: ") + hipGetErrorString(e));
    *out = c;
    return TSMI_OK;
}

void tsmi_destroy(tsmi_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

int tsmi_set_stream(tsmi_ctx *c, void *s) {
    if (!c) return fail(TSMI_E_INVALID, "tsmi_set_stream: null context");
    TSMI_HIP(hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)s;
    return TSMI_OK;
}

int tsmi_set_state(tsmi_ctx *c, const double *re, const double *im) {
    if (c) c->ck_valid = false;
    if (!c || !re || !im) return fail(TSMI_E_INVALID, "tsmi_set_state: null argument");
    int st = upload_plane(c, c->psi_re[c->cur], re);
    if (!st) st = upload_plane(c, c->psi_im[c->cur], im);
    if (st) return st;
    c->has_state = true;
    c->div_pending = false;
    return TSMI_OK;
}

int tsmi_set_state_device(tsmi_ctx *c, const double *re, const double *im) {
    if (c) c->ck_valid = false;
    if (!c || !re || !im) return fail(TSMI_E_INVALID, "tsmi_set_state_device: null argument");
    int st = import_device(c, c->psi_re[c->cur], re);
    if (!st) st = import_device(c, c->psi_im[c->cur], im);
    if (st) return st;
    c->has_state = true;
    c->div_pending = false;
    return TSMI_OK;
}

int tsmi_set_exp_potential_device(tsmi_ctx *c, const double *e_re, const double *e_im) {
    if (c) c->ck_valid = false;
    if (!c || !e_re) return fail(TSMI_E_INVALID, "tsmi_set_exp_potential_device: null argument");
    if (!c->d.imag_time && !e_im) return fail(TSMI_E_INVALID, "tsmi_set_exp_potential_device: real time needs Im(E)");
    const int st = launch_import_exp(c, e_re, c->d.imag_time ? nullptr : e_im);
    if (st) return st;
    if (int we = wrap_exp(c)) return we;
    c->has_e = true;
    return TSMI_OK;
}

int tsmi_set_potential_exp_device(tsmi_ctx *c, const double *v) {
    if (c) c->ck_valid = false;
    if (!c || !v) return fail(TSMI_E_INVALID, "tsmi_set_potential_exp_device: null argument");
    const double dt = c->d.delta_t;
    const int imag = c->d.imag_time ? 1 : 0;
    if (c->esz == 4)
        hipLaunchKernelGGL(exp_from_potential<float>, dim3(2048), dim3(256), 0, c->stream, v, dt, imag, c->corr_re,
                           c->corr_im, (float *)c->e_re, (float *)c->e_im, (float *)c->ph, c->v, c->W, c->H, c->pitch);
    else
        hipLaunchKernelGGL(exp_from_potential<double>, dim3(2048), dim3(256), 0, c->stream, v, dt, imag, c->corr_re,
                           c->corr_im, (double *)c->e_re, (double *)c->e_im, (double *)c->ph, c->v, c->W, c->H,
                           c->pitch);
    TSMI_HIP(hipGetLastError());
    c->ph_mode = true;   // |E| = 1 (real time), E > 0 (imaginary time) by construction
    if (int we = wrap_exp(c)) return we;
    c->has_e = true;
    c->has_v = true;
    return TSMI_OK;
}

int tsmi_set_potential_device(tsmi_ctx *c, const double *v) {
    if (!c || !v) return fail(TSMI_E_INVALID, "tsmi_set_potential_device: null argument");
    TSMI_HIP(hipMemcpy2DAsync(c->v, (size_t)c->pitch * 8, v, (size_t)c->W * 8, (size_t)c->W * 8, c->H,
                              hipMemcpyDeviceToDevice, c->stream));
    c->has_v = true;
    return TSMI_OK;
}

int tsmi_get_state_device(tsmi_ctx *c, double *re, double *im) {
    if (!c || !re || !im) return fail(TSMI_E_INVALID, "tsmi_get_state_device: null argument");
    if (!c->has_state) return fail(TSMI_E_INVALID, "tsmi_get_state_device: no state");
    const double *div = pending_div(c);
    if (div) TSMI_HIP(hipGetLastError());
    const int blocks = 2048;
    for (int k = 0; k < 2; ++k) {
        const void *src = k ? c->psi_im[c->cur] : c->psi_re[c->cur];
        double *dst = k ? im : re;
        if (c->esz == 8)
            hipLaunchKernelGGL(export_plane<double>, dim3(blocks), dim3(256), 0, c->stream, (const double *)src, dst,
                               c->W, c->H, c->pitch, div);
        else
            hipLaunchKernelGGL(export_plane<float>, dim3(blocks), dim3(256), 0, c->stream, (const float *)src, dst,
                               c->W, c->H, c->pitch, div);
        TSMI_HIP(hipGetLastError());
    }
    return TSMI_OK;
}

int tsmi_set_exp_potential(tsmi_ctx *c, const double *e_re, const double *e_im) {
    if (c) c->ck_valid = false;
    if (!c || !e_re) return fail(TSMI_E_INVALID, "tsmi_set_exp_potential: null argument");
    if (!c->d.imag_time && !e_im) return fail(TSMI_E_INVALID, "tsmi_set_exp_potential: real time needs Im(E)");
    // stage the FP64 table on the device; the import kernel applies corr and rounds
    const size_t n = (size_t)c->W * (size_t)c->H;
    const bool cplx = !c->d.imag_time;
    double *tmp = nullptr;
    TSMI_HIP(hipMalloc(&tmp, (cplx ? 2 : 1) * n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, e_re, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && cplx) e = hipMemcpyAsync(tmp + n, e_im, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    int st = TSMI_OK;
    if (e == hipSuccess) st = launch_import_exp(c, tmp, cplx ? tmp + n : nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(TSMI_E_HIP, std::string("exp potential upload: ") + hipGetErrorString(e));
    if (st) return st;
    if (int we = wrap_exp(c)) return we;
    c->has_e = true;
    return TSMI_OK;
}

int tsmi_set_potential(tsmi_ctx *c, const double *v) {
    if (!c || !v) return fail(TSMI_E_INVALID, "tsmi_set_potential: null argument");
    TSMI_HIP(hipMemcpy2DAsync(c->v, (size_t)c->pitch * 8, v, (size_t)c->W * 8, (size_t)c->W * 8, c->H,
                              hipMemcpyHostToDevice, c->stream));
    TSMI_HIP(hipStreamSynchronize(c->stream));
    c->has_v = true;
    return TSMI_OK;
}

int tsmi_evolve(tsmi_ctx *c, int32_t n) {
    if (c) c->ck_valid = false;
    if (!c) return fail(TSMI_E_INVALID, "tsmi_evolve: null context");
    if (n < 0) return fail(TSMI_E_INVALID, "tsmi_evolve: negative step count");
    if (!c->has_state || !c->has_e) return fail(TSMI_E_INVALID, "tsmi_evolve: state and exp potential must be set");
    return evolve_steps(c, n);
}

int tsmi_run_steps(tsmi_ctx *c, int32_t n, int32_t exchange_last) {
    if (c) c->ck_valid = false;
    if (!c) return fail(TSMI_E_INVALID, "tsmi_run_steps: null context");
    if (n < 0) return fail(TSMI_E_INVALID, "tsmi_run_steps: negative step count");
    if (!c->has_state || !c->has_e) return fail(TSMI_E_INVALID, "tsmi_run_steps: state and exp potential must be set");
    return evolve_steps(c, n, exchange_last != 0);
}

int tsmi_step(tsmi_ctx *c, int32_t exchange) {
    if (c) c->ck_valid = false;
    if (!c) return fail(TSMI_E_INVALID, "tsmi_step: null context");
    if (!c->has_state || !c->has_e) return fail(TSMI_E_INVALID, "tsmi_step: state and exp potential must be set");
    return step_once(c, exchange ? 1 : 0);
}

int tsmi_evolve_checkpointed(tsmi_ctx *c, int32_t n, int32_t every) {
    if (!c) return fail(TSMI_E_INVALID, "tsmi_evolve_checkpointed: null context");
    c->ck_valid = false;
    if (n < 0 || every < 1) return fail(TSMI_E_INVALID, "tsmi_evolve_checkpointed: need n >= 0 and every >= 1");
    if (!c->has_state || !c->has_e)
        return fail(TSMI_E_INVALID, "tsmi_evolve_checkpointed: state and exp potential must be set");
    if (!c->adj_ok)
        return fail(TSMI_E_UNSUPPORTED,
                    "tsmi_evolve_checkpointed: the reverse sweep needs real time and, on periodic axes, an inner "
                    "extent of at least twice the halo and no x-periodic tile with a closed y axis");
    const size_t pb = plane_bytes(c);
    const int count = n > 0 ? (n - 1) / every + 1 : 0;
    const size_t need = 2 * (size_t)count * pb;
    if (need > c->ck_cap) {
        if (c->ck) TSMI_HIP(hipFree(c->ck));
        c->ck = nullptr;
        c->ck_cap = 0;
        TSMI_HIP(hipMalloc(&c->ck, need));
        c->ck_cap = need;
    }
    // closed axes, FP32: two steps per pass wherever no checkpoint falls between
    // them (FP64 measured faster at one step per pass on a 1024^2 lattice: 32.7
    // against 34.9 ms for 2000 steps; FP32 37.7 against 30.3)
    const bool pairs = c->sweeps == 2 && !c->halo8 && c->esz == 4;
    for (int i = 0; i < n;) {
        if (i % every == 0) {
            char *dst = (char *)c->ck + (size_t)(i / every) * 2 * pb;
            TSMI_HIP(hipMemcpyAsync(dst, c->psi_base[c->cur][0], pb, hipMemcpyDeviceToDevice, c->stream));
            TSMI_HIP(hipMemcpyAsync(dst + pb, c->psi_base[c->cur][1], pb, hipMemcpyDeviceToDevice, c->stream));
        }
        const bool two = pairs && i + 1 < n && (i + 1) % every != 0;
        // periodic: the halo exchange after every step but the last, as evolve
        const int st = two ? step2_once(c) : step_once(c, i + 1 < n ? 1 : 0);
        if (st) return st;
        i += two ? 2 : 1;
    }
    c->ck_every = every;
    c->ck_steps = n;
    c->ck_count = count;
    c->ck_valid = true;
    return TSMI_OK;
}

int tsmi_adjoint_device(tsmi_ctx *c, const double *g_re, const double *g_im, double *grad_v, double *g0_re,
                        double *g0_im) {
    if (!c || !g_re || !g_im || !grad_v) return fail(TSMI_E_INVALID, "tsmi_adjoint_device: null argument");
    if (!c->ck_valid) return fail(TSMI_E_INVALID, "tsmi_adjoint_device: needs a preceding tsmi_evolve_checkpointed");
    void *l_re = nullptr, *l_im = nullptr;
    int st = reverse_sweep(c, g_re, g_im, &l_re, &l_im);
    if (st) return st;
    TSMI_HIP(hipMemcpy2DAsync(grad_v, (size_t)c->W * 8, c->grad, (size_t)c->pitch * 8, (size_t)c->W * 8, c->H,
                              hipMemcpyDeviceToDevice, c->stream));
    if (g0_re) st = export_device(c, l_re, g0_re);
    if (!st && g0_im) st = export_device(c, l_im, g0_im);
    return st;
}

int tsmi_adjoint(tsmi_ctx *c, const double *g_re, const double *g_im, double *grad_v, double *g0_re, double *g0_im) {
    if (!c || !g_re || !g_im || !grad_v) return fail(TSMI_E_INVALID, "tsmi_adjoint: null argument");
    if (!c->ck_valid) return fail(TSMI_E_INVALID, "tsmi_adjoint: needs a preceding tsmi_evolve_checkpointed");
    const size_t n = (size_t)c->W * (size_t)c->H;
    double *tmp = nullptr;
    TSMI_HIP(hipMalloc(&tmp, 5 * n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, g_re, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tmp + n, g_im, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    int st = e == hipSuccess ? TSMI_OK : fail(TSMI_E_HIP, std::string("tsmi_adjoint upload: ") + hipGetErrorString(e));
    if (!st) st = tsmi_adjoint_device(c, tmp, tmp + n, tmp + 2 * n, tmp + 3 * n, tmp + 4 * n);
    if (!st) {
        e = hipMemcpyAsync(grad_v, tmp + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && g0_re) e = hipMemcpyAsync(g0_re, tmp + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && g0_im) e = hipMemcpyAsync(g0_im, tmp + 4 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = fail(TSMI_E_HIP, std::string("tsmi_adjoint download: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    return st;
}

int tsmi_kernel_timing(tsmi_ctx *c, int32_t enable) {
    if (!c) return fail(TSMI_E_INVALID, "tsmi_kernel_timing: null context");
    c->timing = enable != 0;
    c->ev_used[0] = c->ev_used[1] = 0;
    return TSMI_OK;
}

int tsmi_kernel_time(tsmi_ctx *c, double out[4]) {
    if (!c || !out) return fail(TSMI_E_INVALID, "tsmi_kernel_time: null argument");
    TSMI_HIP(hipStreamSynchronize(c->stream));
    for (int kind = 0; kind < 2; ++kind) {
        double ms = 0.0;
        for (size_t i = 0; i < c->ev_used[kind]; ++i) {
            float t = 0.f;
            TSMI_HIP(hipEventElapsedTime(&t, c->ev[kind][0][i], c->ev[kind][1][i]));
            ms += t;
        }
        out[2 * (1 - kind)] = ms;
        out[2 * (1 - kind) + 1] = (double)c->ev_used[kind];
    }
    return TSMI_OK;
}

int tsmi_squared_norm(tsmi_ctx *c, double *out) {
    if (!c || !out) return fail(TSMI_E_INVALID, "tsmi_squared_norm: null argument");
    if (!c->has_state) return fail(TSMI_E_INVALID, "tsmi_squared_norm: no state");
    double o[kObsOut];
    const int st = observables(c, OBS_KIND_NORM, true, o);
    if (st) return st;
    *out = o[0];
    return TSMI_OK;
}

int tsmi_energy(tsmi_ctx *c, double out[6]) {
    if (!c || !out) return fail(TSMI_E_INVALID, "tsmi_energy: null argument");
    if (!c->has_state || !c->has_v) return fail(TSMI_E_INVALID, "tsmi_energy: state and potential must be set");
    double o[kObsOut];
    // the margin of a one-row lattice ignores the row bounds (src/solver.cpp:356-359)
    const int st = observables(c, OBS_KIND_ENERGY, c->H == 1, o);
    if (st) return st;
    std::memcpy(out, o, 6 * sizeof(double));
    return TSMI_OK;
}

int tsmi_energy_device_noalloc(const tsmi_desc *desc, const void *psi_re, const void *psi_im, int64_t pitch,
                               const double *v, double *out, void *stream) {
    static const char *who = "tsmi_energy_device_noalloc";
    if (!desc || !psi_re || !psi_im || !out) return fail(TSMI_E_INVALID, std::string(who) + ": null argument");
    const tsmi_desc &d = *desc;
    if (int st = check_desc(d, who)) return st;
    const long long W = (long long)d.dim_x + (d.periodic_x ? 2 * d.halo_x : 0);
    const long long H = (long long)d.dim_y + (d.periodic_y ? 2 * d.halo_y : 0);
    if (pitch < W || pitch * H > (1LL << 40)) return fail(TSMI_E_INVALID, std::string(who) + ": bad row pitch");
    // the scratch address on the current device, looked up once per device
    static std::mutex mu;
    static double *ptrs[64] = {};
    int dev = 0;
    TSMI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(TSMI_E_INVALID, std::string(who) + ": device index above 63");
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!ptrs[dev]) {
            void *ptr = nullptr;
            TSMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            ptrs[dev] = (double *)ptr;
        }
        part = ptrs[dev];
    }
    const ObsGeom g = obs_geom(d, (int)W, (int)H, (int)pitch);
    hipStream_t s = (hipStream_t)stream;
    // the margin of a one-row lattice ignores the row bounds (src/solver.cpp:356-359)
    if (d.precision == 32)
        return launch_observables<float>(d, g, (const float *)psi_re, (const float *)psi_im, v, nullptr, nullptr,
                                         OBS_KIND_ENERGY, H == 1, part, nullptr, out, s, who);
    return launch_observables<double>(d, g, (const double *)psi_re, (const double *)psi_im, v, nullptr, nullptr,
                                      OBS_KIND_ENERGY, H == 1, part, nullptr, out, s, who);
}

int tsmi_expected_values(tsmi_ctx *c, double out[10]) {
    if (!c || !out) return fail(TSMI_E_INVALID, "tsmi_expected_values: null argument");
    if (!c->has_state) return fail(TSMI_E_INVALID, "tsmi_expected_values: no state");
    double o[kObsOut];
    const int st = observables(c, OBS_KIND_ALL, false, o);
    if (st) return st;
    std::memcpy(out, o + 6, 10 * sizeof(double));
    return TSMI_OK;
}

int tsmi_get_sample(tsmi_ctx *c, size_t stride, size_t x, size_t y, size_t w, size_t h, double *re, double *im) {
    if (!c || !re || !im) return fail(TSMI_E_INVALID, "tsmi_get_sample: null argument");
    if (!c->has_state) return fail(TSMI_E_INVALID, "tsmi_get_sample: no state");
    if (x + w > (size_t)c->W || y + h > (size_t)c->H || stride < w)
        return fail(TSMI_E_INVALID, "tsmi_get_sample: window outside the tile");
    if (w == 0 || h == 0) return TSMI_OK;
    double div = 1.0;
    int st = current_div(c, &div);
    if (st) return st;
    const size_t es = (size_t)c->esz;
    const size_t off = (y * (size_t)c->pitch + x) * es;
    const bool scale = c->d.imag_time && c->div_pending;
    if (es == 8 && !scale) {
        // FP64 with nothing to apply: straight into the caller's rows
        TSMI_HIP(hipMemcpy2DAsync(re, stride * 8, (const char *)c->psi_re[c->cur] + off, (size_t)c->pitch * 8, w * 8,
                                  h, hipMemcpyDeviceToHost, c->stream));
        TSMI_HIP(hipMemcpy2DAsync(im, stride * 8, (const char *)c->psi_im[c->cur] + off, (size_t)c->pitch * 8, w * 8,
                                  h, hipMemcpyDeviceToHost, c->stream));
        TSMI_HIP(hipStreamSynchronize(c->stream));
        return TSMI_OK;
    }
    // staging (left uninitialised: the copies fill it)
    std::unique_ptr<unsigned char[]> tr(new unsigned char[w * h * es]), ti(new unsigned char[w * h * es]);
    TSMI_HIP(hipMemcpy2DAsync(tr.get(), w * es, (const char *)c->psi_re[c->cur] + off, (size_t)c->pitch * es, w * es,
                              h, hipMemcpyDeviceToHost, c->stream));
    TSMI_HIP(hipMemcpy2DAsync(ti.get(), w * es, (const char *)c->psi_im[c->cur] + off, (size_t)c->pitch * es, w * es,
                              h, hipMemcpyDeviceToHost, c->stream));
    TSMI_HIP(hipStreamSynchronize(c->stream));
    for (size_t j = 0; j < h; ++j)
        for (size_t i = 0; i < w; ++i) {
            double vr, vi;
            if (es == 8) {
                vr = ((const double *)tr.get())[j * w + i];
                vi = ((const double *)ti.get())[j * w + i];
            } else {
                vr = ((const float *)tr.get())[j * w + i];
                vi = ((const float *)ti.get())[j * w + i];
            }
            if (scale) { vr /= div; vi /= div; }   // src/cpukernel.cpp:559-560
            re[j * stride + i] = vr;
            im[j * stride + i] = vi;
        }
    return TSMI_OK;
}

int tsmi_synchronize(tsmi_ctx *c) {
    if (!c) return fail(TSMI_E_INVALID, "tsmi_synchronize: null context");
    TSMI_HIP(hipStreamSynchronize(c->stream));
    return TSMI_OK;
}

int tsmi_geometry(const tsmi_ctx *c, int32_t out[6]) {
    if (!c || !out) return fail(TSMI_E_INVALID, "tsmi_geometry: null argument");
    out[0] = c->W; out[1] = c->H; out[2] = c->pitch;
    out[3] = c->n_strips; out[4] = c->n_segs; out[5] = c->seg_rows;
    return TSMI_OK;
}

int tsmi_plan_geometry(const tsmi_desc *desc, int32_t out[12]) {
    if (!desc || !out) return fail(TSMI_E_INVALID, "tsmi_plan_geometry: null argument");
    const tsmi_desc &d = *desc;
    if (int st = check_desc(d, "tsmi_plan_geometry")) return st;
    if (d.sweeps < 0 || d.sweeps > 2) return fail(TSMI_E_INVALID, "tsmi_plan_geometry: sweeps must be 0, 1 or 2");
    const bool fusable = !d.imag_time && ((d.periodic_x && d.periodic_y) || (!d.periodic_x && !d.periodic_y));
    if (d.sweeps == 2 && !fusable)
        return fail(TSMI_E_UNSUPPORTED,
                    "tsmi_plan_geometry: two steps per pass need real time and both axes closed or both periodic");
    const int sweeps = (d.sweeps != 1 && fusable) ? 2 : 1;
    const bool halo8 = sweeps == 2 && d.periodic_x && d.periodic_y;
    const int W = d.dim_x + (d.periodic_x ? 2 * d.halo_x : 0), H = d.dim_y + (d.periodic_y ? 2 * d.halo_y : 0);
    const int Wd = W + (halo8 ? 2 * d.halo_x : 0), Hd = H + (halo8 ? 2 * d.halo_y : 0);
    const long long pitch = ((long long)Wd + 63) / 64 * 64;
    const int esz = d.precision == 32 ? 4 : 8;
    const int vec = esz == 4 ? (d.imag_time ? kVec32I : kVec32) : 2;
    const int strips = (W + kWave * vec - 2 * kApron - 1) / (kWave * vec - 2 * kApron);
    const int rows = plan_single(d, H, pitch, strips);
    for (int i = 0; i < 12; ++i) out[i] = 0;
    out[0] = W; out[1] = H; out[2] = (int)pitch;
    out[3] = strips; out[4] = (H + rows - 1) / rows; out[5] = rows;
    out[6] = sweeps; out[10] = Wd; out[11] = Hd;
    if (sweeps == 2) {
        // the instance the first pass of a table tabulated from a real
        // potential launches: E as one phase plane, the frame's scaled form
        const StepCoefs k = step_coefs(d, H, esz);
        const bool lhy = d.coupling_lhy != 0.0;
        long simds = 0, resident = 0;
        if (esz == 4) {
            if (k.mode == 0) resident = lhy ? resident_waves(step2_kernel<float, kVec32F, true, 0, true>, &simds)
                                            : resident_waves(step2_kernel<float, kVec32F, false, 0, true>, &simds);
            else if (k.mode == 1) resident = lhy ? resident_waves(step2_kernel<float, kVec32F, true, 1, true>, &simds)
                                                 : resident_waves(step2_kernel<float, kVec32F, false, 1, true>, &simds);
            else resident = lhy ? resident_waves(step2_kernel<float, kVec32F, true, -1, true>, &simds)
                                : resident_waves(step2_kernel<float, kVec32F, false, -1, true>, &simds);
        } else {
            if (k.mode == 0) resident = lhy ? resident_waves(step2_kernel<double, 2, true, 0, true>, &simds)
                                            : resident_waves(step2_kernel<double, 2, false, 0, true>, &simds);
            else if (k.mode == 1) resident = lhy ? resident_waves(step2_kernel<double, 2, true, 1, true>, &simds)
                                                 : resident_waves(step2_kernel<double, 2, false, 1, true>, &simds);
            else resident = lhy ? resident_waves(step2_kernel<double, 2, true, -1, true>, &simds)
                                : resident_waves(step2_kernel<double, 2, false, -1, true>, &simds);
        }
        int s2 = 0, r2 = 0, n2 = 0;
        plan_fused(Wd, Hd, pitch, d.tile_rows, esz == 4 ? kVec32F : 2, resident, simds, &s2, &r2, &n2);
        out[7] = s2; out[8] = n2; out[9] = r2;
    }
    return TSMI_OK;
}

}  // extern "C"
