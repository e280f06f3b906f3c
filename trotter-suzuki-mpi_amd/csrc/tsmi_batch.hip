// tsmi_batch.hip -- C ABI of the batched step (include/tsmi_hip.h, tsmi_batch_*).
//
// No reference counterpart: the reference runs one Solver per lattice.  A batch
// holds n independent single-component lattices (any sizes, each closed or
// periodic on each axis, each with its own dt, mass, spacing and couplings; one
// precision, one time direction and one LHY setting for the whole batch) in one
// device arena, and advances all of them by one persistent launch per step
// (batch_step_kernel) plus, when some lattice is periodic, one exchange launch.
// Each state's step is exactly the single-context step: tsmi_evolve semantics per
// lattice, including the skipped exchange after the last step of a call.
//
// tsmi_batch_evolve records the launch sequence of an n-step call in a HIP graph
// the first time that (n, buffer parity) pair is seen and replays it after.
// tsmi_batch_tune is the autotuner: per size class (lattices of the same tile
// shape) it times the batch step for each candidate segment height and keeps the
// fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "batch_kernel.hpp"
#include "common.hpp"
#include "frame.hpp"
#include "tsmi_hip.h"

using namespace tsmi;

namespace {

constexpr int kBlock = 256;
constexpr int kDefaultSegRows = 64;

#define TSMI_HIPB(call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return set_error(TSMI_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Lattice {
    tsmi_desc d;
    int W = 0, H = 0, pitch = 0;
    int n_strips = 0, seg_rows = 0, n_segs = 0, n_waves = 0, max_waves = 0;
    int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;
    bool fixed_rows = false;
    StepCoefs k;
    size_t psi[2][2] = {{0, 0}, {0, 0}};   // byte offsets in the arena: [buffer][re/im]
    size_t e[2] = {0, 0};
    size_t partials = 0, div = 0, counter = 0;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

struct tsmi_batch {
    int n = 0, esz = 0, vec = 0, mode = 0;
    bool imag = false, lhy = false;
    std::vector<Lattice> lat;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    void *params[2] = {nullptr, nullptr};
    int2 *items = nullptr;
    int n_items = 0, items_cap = 0;
    void *exch[2] = {nullptr, nullptr};
    int n_periodic = 0;
    int cur = 0;
    int grid = 0;
    hipStream_t own = nullptr, stream = nullptr;
    std::map<std::pair<int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

namespace {

void drop_graphs(tsmi_batch *b) {
    for (auto &g : b->graphs) {
        (void)hipGraphExecDestroy(g.second.second);
        (void)hipGraphDestroy(g.second.first);
    }
    b->graphs.clear();
}

void free_batch(tsmi_batch *b) {
    drop_graphs(b);
    void *ptrs[] = {b->arena, b->params[0], b->params[1], b->items, b->exch[0], b->exch[1]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->own) (void)hipStreamDestroy(b->own);
}

void set_rows(Lattice &L, int seg) {
    seg = std::max(2, std::min(seg, std::min(L.H, max_seg_rows(L.pitch))));
    seg = (seg + 1) / 2 * 2;   // segments start on even rows
    L.seg_rows = std::min(seg, L.H + (L.H & 1));
    L.n_segs = (L.H + L.seg_rows - 1) / L.seg_rows;
    L.n_waves = L.n_strips * L.n_segs;
}

template <typename T>
StepParams<T> params_of(const tsmi_batch *b, const Lattice &L, int parity) {
    StepParams<T> p;
    std::memset(&p, 0, sizeof(p));
    char *a = b->arena;
    p.in_re = (const T *)(a + L.psi[parity][0]);
    p.in_im = (const T *)(a + L.psi[parity][1]);
    p.out_re = (T *)(a + L.psi[1 - parity][0]);
    p.out_im = (T *)(a + L.psi[1 - parity][1]);
    p.e_re = (const T *)(a + L.e[0]);
    p.e_im = (const T *)(a + L.e[1]);
    p.W = L.W; p.H = L.H; p.pitch = L.pitch;
    p.n_strips = L.n_strips; p.n_waves = L.n_waves; p.seg_rows = L.seg_rows;
    p.mH = (T)L.k.mH; p.tH = (T)L.k.tH; p.mV = (T)L.k.mV; p.tV = (T)L.k.tV;
    p.unH = (T)L.k.unH; p.unV = (T)L.k.unV;
    p.c_nl = (T)L.k.c_nl_s; p.c_lhy = (T)L.k.c_lhy_s;
    p.ix0 = L.ix0; p.ix1 = L.ix1; p.iy0 = L.iy0; p.iy1 = L.iy1;
    if (b->imag) {
        // the divisor slots follow the buffer parity; a fresh state holds 1.0
        p.div_in = (const double *)(a + L.div) + parity;
        p.div_out = (double *)(a + L.div) + (1 - parity);
        p.partials = (double *)(a + L.partials);
        p.counter = (unsigned *)(a + L.counter);
        p.norm_target = L.d.norm_target;
        p.dxdy = L.d.delta_x * L.d.delta_y;
    }
    return p;
}

template <typename T>
int upload_params(tsmi_batch *b) {
    for (int parity = 0; parity < 2; ++parity) {
        std::vector<StepParams<T>> h(b->n);
        for (int i = 0; i < b->n; ++i) h[i] = params_of<T>(b, b->lat[i], parity);
        TSMI_HIPB(hipMemcpyAsync(b->params[parity], h.data(), h.size() * sizeof(StepParams<T>),
                                 hipMemcpyHostToDevice, b->stream));
        TSMI_HIPB(hipStreamSynchronize(b->stream));   // h is released at the end of the iteration
    }
    return TSMI_OK;
}

// Work items (state, wave) in decreasing order of rows per item, so the longest
// ones are taken first.
int upload_items(tsmi_batch *b) {
    std::vector<int2> items;
    for (int i = 0; i < b->n; ++i)
        for (int w = 0; w < b->lat[i].n_waves; ++w) items.push_back(make_int2(i, w));
    std::stable_sort(items.begin(), items.end(), [&](const int2 &x, const int2 &y) {
        return b->lat[x.x].seg_rows > b->lat[y.x].seg_rows;
    });
    if ((int)items.size() > b->items_cap) {
        if (b->items) TSMI_HIPB(hipFree(b->items));
        b->items = nullptr;
        b->items_cap = 0;
        TSMI_HIPB(hipMalloc((void **)&b->items, items.size() * sizeof(int2)));
        b->items_cap = (int)items.size();
    }
    TSMI_HIPB(hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
    b->n_items = (int)items.size();
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

int upload_layout(tsmi_batch *b) {
    const int st = b->esz == 4 ? upload_params<float>(b) : upload_params<double>(b);
    if (st) return st;
    drop_graphs(b);
    return upload_items(b);
}

template <typename T>
int upload_exchange(tsmi_batch *b) {
    for (int parity = 0; parity < 2; ++parity) {
        std::vector<ExchParams<T>> h;
        for (const Lattice &L : b->lat) {
            if (!L.d.periodic_x && !L.d.periodic_y) continue;
            ExchParams<T> q;
            // after a step of parity p the state lives in buffer 1 - p
            q.re = (T *)(b->arena + L.psi[1 - parity][0]);
            q.im = (T *)(b->arena + L.psi[1 - parity][1]);
            q.W = L.W; q.H = L.H; q.pitch = L.pitch;
            q.px = L.d.periodic_x; q.py = L.d.periodic_y; q.hx = L.d.halo_x; q.hy = L.d.halo_y;
            h.push_back(q);
        }
        b->n_periodic = (int)h.size();
        if (h.empty()) continue;
        TSMI_HIPB(hipMalloc(&b->exch[parity], h.size() * sizeof(ExchParams<T>)));
        TSMI_HIPB(hipMemcpyAsync(b->exch[parity], h.data(), h.size() * sizeof(ExchParams<T>), hipMemcpyHostToDevice,
                                 b->stream));
        TSMI_HIPB(hipStreamSynchronize(b->stream));
    }
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY, int MODE>
int launch_batch_step(tsmi_batch *b, int parity) {
    hipLaunchKernelGGL((batch_step_kernel<T, IMAG, VEC, LHY, MODE>), dim3(b->grid), dim3(kBlock), 0, b->stream,
                       (const StepParams<T> *)b->params[parity], (const int2 *)b->items, b->n_items);
    TSMI_HIPB(hipGetLastError());
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY>
int batch_step_mode(tsmi_batch *b, int parity) {
    if constexpr (IMAG) {
        return launch_batch_step<T, IMAG, VEC, LHY, 0>(b, parity);
    } else {
        if (b->mode == 0) return launch_batch_step<T, IMAG, VEC, LHY, 0>(b, parity);
        if (b->mode == 1) return launch_batch_step<T, IMAG, VEC, LHY, 1>(b, parity);
        return launch_batch_step<T, IMAG, VEC, LHY, -1>(b, parity);
    }
}

template <typename T, int VEC>
int batch_step_typed(tsmi_batch *b, int parity) {
    if (b->imag) return b->lhy ? batch_step_mode<T, true, VEC, true>(b, parity) : batch_step_mode<T, true, VEC, false>(b, parity);
    return b->lhy ? batch_step_mode<T, false, VEC, true>(b, parity) : batch_step_mode<T, false, VEC, false>(b, parity);
}

// One step of every state from buffer `parity`; the exchange when asked.
int batch_step(tsmi_batch *b, int parity, bool exchange) {
    int st = b->esz == 4 ? batch_step_typed<float, 4>(b, parity) : batch_step_typed<double, 2>(b, parity);
    if (st) return st;
    if (exchange && b->n_periodic > 0) {
        if (b->esz == 4)
            hipLaunchKernelGGL(batch_exchange_kernel<float>, dim3(b->n_periodic), dim3(kBlock), 0, b->stream,
                               (const ExchParams<float> *)b->exch[parity]);
        else
            hipLaunchKernelGGL(batch_exchange_kernel<double>, dim3(b->n_periodic), dim3(kBlock), 0, b->stream,
                               (const ExchParams<double> *)b->exch[parity]);
        TSMI_HIPB(hipGetLastError());
    }
    return TSMI_OK;
}

// n steps from buffer `parity`, eagerly.
int batch_run(tsmi_batch *b, int n, int parity) {
    for (int i = 0; i < n; ++i) {
        const int st = batch_step(b, (parity + i) & 1, i != n - 1);
        if (st) return st;
    }
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY>
hipError_t batch_occupancy(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, batch_step_kernel<T, IMAG, VEC, LHY, 0>, kBlock, 0);
}

hipError_t batch_occupancy_of(const tsmi_batch *b, int *blocks) {
    if (b->esz == 4) {
        if (b->imag) return b->lhy ? batch_occupancy<float, true, 4, true>(blocks) : batch_occupancy<float, true, 4, false>(blocks);
        return b->lhy ? batch_occupancy<float, false, 4, true>(blocks) : batch_occupancy<float, false, 4, false>(blocks);
    }
    if (b->imag) return b->lhy ? batch_occupancy<double, true, 2, true>(blocks) : batch_occupancy<double, true, 2, false>(blocks);
    return b->lhy ? batch_occupancy<double, false, 2, true>(blocks) : batch_occupancy<double, false, 2, false>(blocks);
}

Lattice *lattice_at(tsmi_batch *b, int i, const char *who) {
    if (!b || i < 0 || i >= b->n) {
        set_error(TSMI_E_INVALID, std::string(who) + ": no such lattice");
        return nullptr;
    }
    return &b->lat[i];
}

// W x H FP64 host plane -> pitched plane of the batch precision (rounded once)
int upload_host_plane(tsmi_batch *b, const Lattice &L, size_t off, const double *src) {
    const size_t n = (size_t)L.W * (size_t)L.H;
    char *dst = b->arena + off;
    if (b->esz == 8) {
        TSMI_HIPB(hipMemcpy2DAsync(dst, (size_t)L.pitch * 8, src, (size_t)L.W * 8, (size_t)L.W * 8, L.H,
                                   hipMemcpyHostToDevice, b->stream));
        TSMI_HIPB(hipStreamSynchronize(b->stream));
    } else {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
        TSMI_HIPB(hipMemcpy2DAsync(dst, (size_t)L.pitch * 4, tmp.data(), (size_t)L.W * 4, (size_t)L.W * 4, L.H,
                                   hipMemcpyHostToDevice, b->stream));
        TSMI_HIPB(hipStreamSynchronize(b->stream));
    }
    return TSMI_OK;
}

int reset_divisor(tsmi_batch *b, const Lattice &L) {
    if (!b->imag) return TSMI_OK;
    const double one[2] = {1.0, 1.0};
    TSMI_HIPB(hipMemcpyAsync(b->arena + L.div, one, sizeof(one), hipMemcpyHostToDevice, b->stream));
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

}  // namespace

extern "C" {

int tsmi_batch_create(const tsmi_desc *descs, int32_t n, tsmi_batch **out) {
    if (!descs || !out || n < 1) return set_error(TSMI_E_INVALID, "tsmi_batch_create: need n >= 1 descriptors");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (int st = check_desc(descs[i], "tsmi_batch_create")) return st;
        if (descs[i].precision != descs[0].precision || (descs[i].imag_time != 0) != (descs[0].imag_time != 0) ||
            (descs[i].coupling_lhy != 0.0) != (descs[0].coupling_lhy != 0.0))
            return set_error(TSMI_E_UNSUPPORTED,
                             "tsmi_batch_create: all lattices need the same precision, time direction and LHY setting");
    }
    tsmi_batch *b = new tsmi_batch();
    b->n = n;
    b->esz = descs[0].precision == 32 ? 4 : 8;
    b->vec = b->esz == 4 ? 4 : 2;
    b->imag = descs[0].imag_time != 0;
    b->lhy = descs[0].coupling_lhy != 0.0;
    b->lat.resize(n);
    const int sw = kWave * b->vec - 2 * kApron;
    size_t off = 0;
    int mode = -2;
    for (int i = 0; i < n; ++i) {
        Lattice &L = b->lat[i];
        const tsmi_desc &d = descs[i];
        L.d = d;
        L.W = d.dim_x + (d.periodic_x ? 2 * d.halo_x : 0);
        L.H = d.dim_y + (d.periodic_y ? 2 * d.halo_y : 0);
        L.pitch = (L.W + 63) / 64 * 64;
        L.n_strips = (L.W + sw - 1) / sw;
        L.fixed_rows = d.tile_rows > 0;
        set_rows(L, d.tile_rows > 0 ? d.tile_rows : kDefaultSegRows);
        L.max_waves = L.n_strips * ((L.H + 1) / 2);   // the smallest segment height, 2 rows
        L.ix0 = d.periodic_x ? d.halo_x : 0; L.ix1 = L.W - L.ix0;
        L.iy0 = d.periodic_y ? d.halo_y : 0; L.iy1 = L.H - L.iy0;
        L.k = step_coefs(d, L.H, b->esz);
        mode = mode == -2 ? L.k.mode : (mode == L.k.mode ? mode : -1);
        const size_t pb = align256((size_t)L.pitch * (size_t)L.H * (size_t)b->esz);
        for (int buf = 0; buf < 2; ++buf)
            for (int c = 0; c < 2; ++c) { L.psi[buf][c] = off; off += pb; }
        L.e[0] = off; off += pb;
        L.e[1] = off; off += pb;
        L.partials = off; off += align256((size_t)L.max_waves * sizeof(double));
        L.div = off; off += 256;
        L.counter = off; off += 256;
    }
    b->mode = mode;
    b->arena_bytes = off;
    auto bail = [&](int code) {
        free_batch(b);
        delete b;
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    b->stream = b->own;
    e = hipMalloc((void **)&b->arena, b->arena_bytes);
    if (e != hipSuccess) return bail(set_error(TSMI_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e)));
    if (e == hipSuccess) e = hipMemsetAsync(b->arena, 0, b->arena_bytes, b->stream);
    const size_t psz = b->esz == 4 ? sizeof(StepParams<float>) : sizeof(StepParams<double>);
    if (e == hipSuccess) e = hipMalloc(&b->params[0], (size_t)n * psz);
    if (e == hipSuccess) e = hipMalloc(&b->params[1], (size_t)n * psz);
    if (e != hipSuccess) return bail(set_error(TSMI_E_OOM, std::string("batch allocation: ") + hipGetErrorString(e)));
    for (const Lattice &L : b->lat)
        if (int st = reset_divisor(b, L)) return bail(st);
    int st = b->esz == 4 ? upload_exchange<float>(b) : upload_exchange<double>(b);
    if (!st) st = upload_layout(b);
    if (st) return bail(st);
    int dev = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = batch_occupancy_of(b, &per_cu);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e)));
    b->grid = std::max(1, cus * std::max(1, per_cu));
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("batch set-up: ") + hipGetErrorString(e)));
    *out = b;
    return TSMI_OK;
}

void tsmi_batch_destroy(tsmi_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
}

int tsmi_batch_set_state(tsmi_batch *b, int32_t i, const double *re, const double *im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_set_state");
    if (!L) return TSMI_E_INVALID;
    if (!re || !im) return set_error(TSMI_E_INVALID, "tsmi_batch_set_state: null argument");
    int st = upload_host_plane(b, *L, L->psi[b->cur][0], re);
    if (!st) st = upload_host_plane(b, *L, L->psi[b->cur][1], im);
    if (!st) st = reset_divisor(b, *L);
    return st;
}

int tsmi_batch_set_exp_potential(tsmi_batch *b, int32_t i, const double *e_re, const double *e_im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_set_exp_potential");
    if (!L) return TSMI_E_INVALID;
    if (!e_re || (!b->imag && !e_im)) return set_error(TSMI_E_INVALID, "tsmi_batch_set_exp_potential: null argument");
    // E times the frame correction, in FP64, rounded once
    const size_t n = (size_t)L->W * (size_t)L->H;
    std::vector<double> re(n), im(n, 0.0);
    for (size_t k = 0; k < n; ++k) {
        const double er = e_re[k], ei = b->imag ? 0.0 : e_im[k];
        re[k] = er * L->k.corr_re - ei * L->k.corr_im;
        im[k] = er * L->k.corr_im + ei * L->k.corr_re;
    }
    int st = upload_host_plane(b, *L, L->e[0], re.data());
    if (!st) st = upload_host_plane(b, *L, L->e[1], im.data());
    return st;
}

int tsmi_batch_evolve(tsmi_batch *b, int32_t n) {
    if (!b) return set_error(TSMI_E_INVALID, "tsmi_batch_evolve: null batch");
    if (n < 0) return set_error(TSMI_E_INVALID, "tsmi_batch_evolve: negative step count");
    if (n == 0) return TSMI_OK;
    const std::pair<int, int> key(n, b->cur);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        TSMI_HIPB(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int st = batch_run(b, n, b->cur);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &captured);
        if (st) {
            if (captured) (void)hipGraphDestroy(captured);
            return st;
        }
        if (e != hipSuccess) return set_error(TSMI_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        g = captured;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(TSMI_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        it = b->graphs.emplace(key, std::make_pair(g, x)).first;
    }
    TSMI_HIPB(hipGraphLaunch(it->second.second, b->stream));
    b->cur = (b->cur + n) & 1;
    return TSMI_OK;
}

int tsmi_batch_get_state(tsmi_batch *b, int32_t i, double *re, double *im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_get_state");
    if (!L) return TSMI_E_INVALID;
    if (!re || !im) return set_error(TSMI_E_INVALID, "tsmi_batch_get_state: null argument");
    const size_t n = (size_t)L->W * (size_t)L->H, es = (size_t)b->esz;
    std::vector<unsigned char> tr(n * es), ti(n * es);
    double div = 1.0;
    TSMI_HIPB(hipMemcpy2DAsync(tr.data(), (size_t)L->W * es, b->arena + L->psi[b->cur][0], (size_t)L->pitch * es,
                               (size_t)L->W * es, L->H, hipMemcpyDeviceToHost, b->stream));
    TSMI_HIPB(hipMemcpy2DAsync(ti.data(), (size_t)L->W * es, b->arena + L->psi[b->cur][1], (size_t)L->pitch * es,
                               (size_t)L->W * es, L->H, hipMemcpyDeviceToHost, b->stream));
    if (b->imag)
        TSMI_HIPB(hipMemcpyAsync(&div, (const double *)(b->arena + L->div) + b->cur, sizeof(double),
                                 hipMemcpyDeviceToHost, b->stream));
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    for (size_t k = 0; k < n; ++k) {
        double vr, vi;
        if (es == 8) { vr = ((const double *)tr.data())[k]; vi = ((const double *)ti.data())[k]; }
        else         { vr = ((const float *)tr.data())[k];  vi = ((const float *)ti.data())[k]; }
        re[k] = vr / div;
        im[k] = vi / div;
    }
    return TSMI_OK;
}

int tsmi_batch_tune(tsmi_batch *b, int32_t steps, double *ms_per_step) {
    if (!b || steps < 1) return set_error(TSMI_E_INVALID, "tsmi_batch_tune: need a batch and steps >= 1");
    // the trials advance the states: keep a copy of every buffer and restore it
    char *backup = nullptr;
    TSMI_HIPB(hipMalloc((void **)&backup, b->arena_bytes));
    auto restore = [&]() {
        (void)hipMemcpyAsync(b->arena, backup, b->arena_bytes, hipMemcpyDeviceToDevice, b->stream);
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(backup);
    };
    hipError_t e = hipMemcpyAsync(backup, b->arena, b->arena_bytes, hipMemcpyDeviceToDevice, b->stream);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e != hipSuccess) {
        restore();
        return set_error(TSMI_E_HIP, std::string("tsmi_batch_tune: ") + hipGetErrorString(e));
    }
    int parity = b->cur;
    auto trial = [&](float *ms) -> int {
        if (int st = upload_layout(b)) return st;
        if (int st = batch_run(b, 1, parity)) return st;   // warm-up
        parity ^= 1;
        TSMI_HIPB(hipEventRecord(t0, b->stream));
        if (int st = batch_run(b, steps, parity)) return st;
        TSMI_HIPB(hipEventRecord(t1, b->stream));
        TSMI_HIPB(hipEventSynchronize(t1));
        parity = (parity + steps) & 1;
        TSMI_HIPB(hipEventElapsedTime(ms, t0, t1));
        return TSMI_OK;
    };
    // size classes: lattices of the same tile shape, largest first
    std::map<std::pair<int, int>, std::vector<int>> classes;
    for (int i = 0; i < b->n; ++i)
        if (!b->lat[i].fixed_rows) classes[{b->lat[i].H, b->lat[i].W}].push_back(i);
    std::vector<std::pair<int, int>> order;
    for (auto &c : classes) order.push_back(c.first);
    std::sort(order.begin(), order.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
        return (long)x.first * x.second > (long)y.first * y.second;
    });
    static const int cand[] = {8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
    int st = TSMI_OK;
    float best_all = 0.f;
    for (const auto &key : order) {
        const std::vector<int> &members = classes[key];
        const int keep = b->lat[members[0]].seg_rows;
        int best_rows = keep;
        float best = -1.f;
        for (int rows : cand) {
            if (rows > key.first + 1 && rows != cand[0]) break;
            for (int i : members) set_rows(b->lat[i], rows);
            float ms = 0.f;
            st = trial(&ms);
            if (st) break;
            if (best < 0.f || ms < best) { best = ms; best_rows = rows; }
        }
        if (st) break;
        for (int i : members) set_rows(b->lat[i], best_rows);
        best_all = best;
    }
    if (!st) st = upload_layout(b);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    restore();
    if (st) return st;
    if (ms_per_step) *ms_per_step = best_all / steps;
    return TSMI_OK;
}

int tsmi_batch_info(const tsmi_batch *b, int32_t i, int32_t out[6]) {
    if (!b || !out || i < 0 || i >= b->n) return set_error(TSMI_E_INVALID, "tsmi_batch_info: no such lattice");
    const Lattice &L = b->lat[i];
    out[0] = L.W; out[1] = L.H; out[2] = L.pitch;
    out[3] = L.n_strips; out[4] = L.n_segs; out[5] = L.seg_rows;
    return TSMI_OK;
}

int tsmi_batch_synchronize(tsmi_batch *b) {
    if (!b) return set_error(TSMI_E_INVALID, "tsmi_batch_synchronize: null batch");
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

}  // extern "C"
