// tsmi_batch.hip -- C ABI of the batched step (include/tsmi_hip.h, tsmi_batch_*).
//
// No reference counterpart: the reference runs one Solver per lattice.  A batch
// holds n independent single-component lattices (any sizes, each closed or
// periodic on each axis, each with its own dt, mass, spacing and couplings; one
// precision, one time direction and one LHY setting for the whole batch) in one
// device arena, and advances all of them with persistent launches that walk
// (lattice, wave) work items.  Each state's evolution is exactly the
// single-context one: tsmi_evolve semantics per lattice, including the skipped
// exchange after the last step of a call.
//
// Lattices are in one of two groups.  Group 1 holds the lattices that take two
// real-time steps per pass (both axes closed or both periodic, sweeps != 1);
// group 0 the rest (imaginary time, mixed boundaries, sweeps == 1), stepped one
// step per pass.  Group 1 follows the single context's sequence: on the 8-wide
// halo layout of a doubly periodic lattice, a single step, then fused passes
// each followed by the 8-wide wrap, then the last single step.  The two groups
// are independent; when both are present their launches go to two streams.
//
// tsmi_batch_evolve records the launch sequence of an n-step call in a HIP graph
// the first time that (n, buffer parities) is seen and replays it after.
// tsmi_batch_tune is the autotuner: per size class (lattices of the same group
// and tile shape) it times the batch for each candidate segment height of that
// class's dominant pass and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
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
constexpr int kVecF = 2;                       // elements per lane of the fused passes
constexpr int kExchChunks = 64;                // most workgroups per lattice of an exchange launch

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
    // two steps per pass: the fused decomposition over the Wd x Hd allocation
    // (the reference tile plus 4 more halo cells per periodic side when halo8).
    // fusable: the layout allows both passes; fused: the pass in use; tunable:
    // the autotuner may switch between them (descriptor sweeps 0)
    bool fusable = false, fused = false, tunable = false, halo8 = false;
    int Wd = 0, Hd = 0;
    int n_strips2 = 0, seg_rows2 = 0, n_segs2 = 0, n_waves2 = 0;
    StepCoefs k;
    size_t shift = 0;                         // bytes from an allocation to its reference-tile view
    size_t psi[2][2] = {{0, 0}, {0, 0}};      // arena offsets of the views: [buffer][re/im]
    size_t e[2] = {0, 0};                     // view of E
    size_t e8[2] = {0, 0};                    // allocation of the E the fused passes read
    size_t ph = 0, ph8 = 0;                   // the one-plane form of E: view, and what the fused passes read
    bool ph_ok = false;                       // E has that form (|E| = 1, or E > 0)
    // imaginary time: partial sums of |psi|^2 per buffer parity (max_waves
    // each), their two counts, and a divisor for host-side readers
    size_t partials = 0, div = 0, part_n = 0;
    int group() const { return fused ? 1 : 0; }
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

struct GraphRec {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    int end[2] = {0, 0};
};

}  // namespace

struct tsmi_batch {
    int n = 0, esz = 0, vec = 0, mode = 0;
    bool imag = false, lhy = false;
    std::vector<Lattice> lat;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    void *params[2] = {nullptr, nullptr};     // single-step parameters per buffer parity
    void *params2[2] = {nullptr, nullptr};    // fused-pass parameters per buffer parity
    // work items: [0] single steps of group 0, [1] single steps of group 1,
    // [2] fused passes of group 1
    int2 *items[3] = {nullptr, nullptr, nullptr};
    int n_items[3] = {0, 0, 0}, items_cap[3] = {0, 0, 0};
    void *exch[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [group][buffer holding the state]
    int n_exch[2] = {0, 0}, exch_chunks[2] = {1, 1};
    int cur[2] = {0, 0};                      // buffer holding the states, per group
    bool ph = false;                          // every lattice's E has the one-plane form
    int grid[2] = {0, 0}, grid2[2] = {0, 0};   // persistent grids of the E form [0] and the phase form [1]
    hipStream_t own = nullptr, stream = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::map<std::tuple<int, int, int>, GraphRec> graphs;
};

namespace {

void drop_graphs(tsmi_batch *b) {
    for (auto &g : b->graphs) {
        (void)hipGraphExecDestroy(g.second.x);
        (void)hipGraphDestroy(g.second.g);
    }
    b->graphs.clear();
}

void free_batch(tsmi_batch *b) {
    drop_graphs(b);
    void *ptrs[] = {b->arena,    b->params[0], b->params[1], b->params2[0], b->params2[1], b->items[0], b->items[1],
                    b->items[2], b->exch[0][0], b->exch[0][1], b->exch[1][0], b->exch[1][1]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->own) (void)hipStreamDestroy(b->own);
}

void set_rows(Lattice &L, int seg) {
    seg = std::max(2, std::min(seg, std::min(L.H, max_seg_rows(L.pitch, 2 * kApron))));
    seg = (seg + 1) / 2 * 2;   // segments start on even rows
    L.seg_rows = std::min(seg, L.H + (L.H & 1));
    L.n_segs = (L.H + L.seg_rows - 1) / L.seg_rows;
    L.n_waves = L.n_strips * L.n_segs;
}

void set_rows2(Lattice &L, int seg) {
    if (!L.fusable) return;
    seg = std::max(2, std::min(seg, std::min(L.Hd, max_seg_rows(L.pitch, 4 * kApron))));
    seg = (seg + 1) / 2 * 2;
    L.seg_rows2 = std::min(seg, L.Hd + (L.Hd & 1));
    L.n_segs2 = (L.Hd + L.seg_rows2 - 1) / L.seg_rows2;
    L.n_waves2 = L.n_strips2 * L.n_segs2;
}

template <typename T>
void set_coefs(StepParams<T> &p, const Lattice &L) {
    p.mH = (T)L.k.mH; p.tH = (T)L.k.tH; p.mV = (T)L.k.mV; p.tV = (T)L.k.tV;
    p.unH = (T)L.k.unH; p.unV = (T)L.k.unV;
    p.c_nl = (T)L.k.c_nl_s; p.c_lhy = (T)L.k.c_lhy_s;
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
    p.ph = (const T *)(a + L.ph);
    p.mag = b->imag ? 1.0 : L.k.corr_abs;
    p.W = L.W; p.H = L.H; p.pitch = L.pitch;
    p.n_strips = L.n_strips; p.n_waves = L.n_waves; p.seg_rows = L.seg_rows;
    set_coefs(p, L);
    p.ix0 = L.ix0; p.ix1 = L.ix1; p.iy0 = L.iy0; p.iy1 = L.iy1;
    if (b->imag) {
        // the partial-sum slots follow the buffer parity; a fresh state has a
        // count of 0 (divisor 1), and the counts travel with the sums, so a
        // pending renormalisation survives a change of segment heights
        const size_t mw = (size_t)L.max_waves;
        p.part_in = (const double *)(a + L.partials) + (size_t)parity * mw;
        p.n_in = (const int *)(a + L.part_n) + parity;
        p.part_out = (double *)(a + L.partials) + (size_t)(1 - parity) * mw;
        p.n_out = (int *)(a + L.part_n) + (1 - parity);
        p.norm_target = L.d.norm_target;
        p.dxdy = L.d.delta_x * L.d.delta_y;
    }
    return p;
}

// A fused pass reads and writes whole allocations (fused_kernel.hpp).
template <typename T>
StepParams<T> params2_of(const tsmi_batch *b, const Lattice &L, int parity) {
    StepParams<T> p;
    std::memset(&p, 0, sizeof(p));
    if (!L.fused) return p;
    char *a = b->arena;
    p.in_re = (const T *)(a + L.psi[parity][0] - L.shift);
    p.in_im = (const T *)(a + L.psi[parity][1] - L.shift);
    p.out_re = (T *)(a + L.psi[1 - parity][0] - L.shift);
    p.out_im = (T *)(a + L.psi[1 - parity][1] - L.shift);
    p.e_re = (const T *)(a + L.e8[0]);
    p.e_im = (const T *)(a + L.e8[1]);
    p.ph = (const T *)(a + L.ph8);
    p.mag = L.k.corr_abs;
    p.W = L.Wd; p.H = L.Hd; p.pitch = L.pitch;
    p.n_strips = L.n_strips2; p.n_waves = L.n_waves2; p.seg_rows = L.seg_rows2;
    set_coefs(p, L);
    return p;
}

template <typename T>
int upload_params(tsmi_batch *b) {
    for (int parity = 0; parity < 2; ++parity) {
        std::vector<StepParams<T>> h(b->n), h2(b->n);
        for (int i = 0; i < b->n; ++i) {
            h[i] = params_of<T>(b, b->lat[i], parity);
            h2[i] = params2_of<T>(b, b->lat[i], parity);
        }
        TSMI_HIPB(hipMemcpyAsync(b->params[parity], h.data(), h.size() * sizeof(StepParams<T>),
                                 hipMemcpyHostToDevice, b->stream));
        TSMI_HIPB(hipMemcpyAsync(b->params2[parity], h2.data(), h2.size() * sizeof(StepParams<T>),
                                 hipMemcpyHostToDevice, b->stream));
        TSMI_HIPB(hipStreamSynchronize(b->stream));   // h, h2 are released at the end of the iteration
    }
    return TSMI_OK;
}

// Work items (lattice, wave) of one list in decreasing order of rows per item,
// so the longest ones are taken first.
int upload_item_list(tsmi_batch *b, int list) {
    std::vector<int2> items;
    std::vector<int> rows;
    for (int i = 0; i < b->n; ++i) {
        const Lattice &L = b->lat[i];
        if ((list == 0 && L.fused) || (list > 0 && !L.fused)) continue;
        const int waves = list == 2 ? L.n_waves2 : L.n_waves;
        for (int w = 0; w < waves; ++w) items.push_back(make_int2(i, w));
    }
    auto rows_of = [&](const int2 &x) { return list == 2 ? b->lat[x.x].seg_rows2 : b->lat[x.x].seg_rows; };
    std::stable_sort(items.begin(), items.end(), [&](const int2 &x, const int2 &y) { return rows_of(x) > rows_of(y); });
    b->n_items[list] = (int)items.size();
    if (items.empty()) return TSMI_OK;
    if ((int)items.size() > b->items_cap[list]) {
        if (b->items[list]) TSMI_HIPB(hipFree(b->items[list]));
        b->items[list] = nullptr;
        b->items_cap[list] = 0;
        TSMI_HIPB(hipMalloc((void **)&b->items[list], items.size() * sizeof(int2)));
        b->items_cap[list] = (int)items.size();
    }
    TSMI_HIPB(hipMemcpyAsync(b->items[list], items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice,
                             b->stream));
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

int upload_layout(tsmi_batch *b) {
    const int st = b->esz == 4 ? upload_params<float>(b) : upload_params<double>(b);
    if (st) return st;
    drop_graphs(b);
    for (int list = 0; list < 3; ++list)
        if (int s = upload_item_list(b, list)) return s;
    return TSMI_OK;
}

// Halo exchange lists per group and per buffer: the reference tile's exchange
// for group 0, the 8-wide wrap over the allocation for the doubly periodic
// lattices of group 1 (whose inner 4 cells are what the 4-wide exchange writes).
template <typename T>
int upload_exchange(tsmi_batch *b) {
    for (int g = 0; g < 2; ++g) {
        for (int buf = 0; buf < 2; ++buf) {
            std::vector<ExchParams<T>> h;
            long most = 0;
            for (const Lattice &L : b->lat) {
                if (L.group() != g || (!L.d.periodic_x && !L.d.periodic_y)) continue;
                ExchParams<T> q;
                if (L.halo8 && L.fused) {
                    q.re = (T *)(b->arena + L.psi[buf][0] - L.shift);
                    q.im = (T *)(b->arena + L.psi[buf][1] - L.shift);
                    q.W = L.Wd; q.H = L.Hd; q.hx = 2 * L.d.halo_x; q.hy = 2 * L.d.halo_y;
                } else {
                    q.re = (T *)(b->arena + L.psi[buf][0]);
                    q.im = (T *)(b->arena + L.psi[buf][1]);
                    q.W = L.W; q.H = L.H; q.hx = L.d.halo_x; q.hy = L.d.halo_y;
                }
                q.pitch = L.pitch;
                q.px = L.d.periodic_x; q.py = L.d.periodic_y;
                most = std::max(most, (long)exchange_cells(q.W, q.H, q.px, q.py, q.hx, q.hy));
                h.push_back(q);
            }
            b->n_exch[g] = (int)h.size();
            b->exch_chunks[g] = (int)std::min<long>(kExchChunks, std::max<long>(1, (most + 1023) / 1024));
            if (b->exch[g][buf]) {
                TSMI_HIPB(hipStreamSynchronize(b->stream));
                TSMI_HIPB(hipFree(b->exch[g][buf]));
                b->exch[g][buf] = nullptr;
            }
            if (h.empty()) continue;
            TSMI_HIPB(hipMalloc(&b->exch[g][buf], h.size() * sizeof(ExchParams<T>)));
            TSMI_HIPB(hipMemcpyAsync(b->exch[g][buf], h.data(), h.size() * sizeof(ExchParams<T>),
                                     hipMemcpyHostToDevice, b->stream));
            TSMI_HIPB(hipStreamSynchronize(b->stream));
        }
    }
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY, int MODE>
int launch_single(tsmi_batch *b, hipStream_t s, int list, int parity) {
    if (b->ph)
        hipLaunchKernelGGL((batch_step_kernel<T, IMAG, VEC, LHY, MODE, true>), dim3(b->grid[1]), dim3(kBlock), 0, s,
                           (const StepParams<T> *)b->params[parity], (const int2 *)b->items[list], b->n_items[list]);
    else
        hipLaunchKernelGGL((batch_step_kernel<T, IMAG, VEC, LHY, MODE, false>), dim3(b->grid[0]), dim3(kBlock), 0, s,
                           (const StepParams<T> *)b->params[parity], (const int2 *)b->items[list], b->n_items[list]);
    TSMI_HIPB(hipGetLastError());
    return TSMI_OK;
}

template <typename T, bool IMAG, int VEC, bool LHY>
int single_mode(tsmi_batch *b, hipStream_t s, int list, int parity) {
    if constexpr (IMAG) {
        return launch_single<T, IMAG, VEC, LHY, 0>(b, s, list, parity);
    } else {
        if (b->mode == 0) return launch_single<T, IMAG, VEC, LHY, 0>(b, s, list, parity);
        if (b->mode == 1) return launch_single<T, IMAG, VEC, LHY, 1>(b, s, list, parity);
        return launch_single<T, IMAG, VEC, LHY, -1>(b, s, list, parity);
    }
}

template <typename T, int VEC>
int single_typed(tsmi_batch *b, hipStream_t s, int list, int parity) {
    if (b->imag)
        return b->lhy ? single_mode<T, true, VEC, true>(b, s, list, parity)
                      : single_mode<T, true, VEC, false>(b, s, list, parity);
    return b->lhy ? single_mode<T, false, VEC, true>(b, s, list, parity)
                  : single_mode<T, false, VEC, false>(b, s, list, parity);
}

template <typename T, bool LHY, int MODE>
int launch_fused(tsmi_batch *b, hipStream_t s, int parity) {
    if (b->ph)
        hipLaunchKernelGGL((batch_step2_kernel<T, kVecF, LHY, MODE, true>), dim3(b->grid2[1]), dim3(kBlock), 0, s,
                           (const StepParams<T> *)b->params2[parity], (const int2 *)b->items[2], b->n_items[2]);
    else
        hipLaunchKernelGGL((batch_step2_kernel<T, kVecF, LHY, MODE, false>), dim3(b->grid2[0]), dim3(kBlock), 0, s,
                           (const StepParams<T> *)b->params2[parity], (const int2 *)b->items[2], b->n_items[2]);
    TSMI_HIPB(hipGetLastError());
    return TSMI_OK;
}

template <typename T, bool LHY>
int fused_mode(tsmi_batch *b, hipStream_t s, int parity) {
    if (b->mode == 0) return launch_fused<T, LHY, 0>(b, s, parity);
    if (b->mode == 1) return launch_fused<T, LHY, 1>(b, s, parity);
    return launch_fused<T, LHY, -1>(b, s, parity);
}

int exchange(tsmi_batch *b, hipStream_t s, int g, int buf) {
    if (b->n_exch[g] == 0) return TSMI_OK;
    const dim3 grid(b->n_exch[g], b->exch_chunks[g]);
    if (b->esz == 4)
        hipLaunchKernelGGL(batch_exchange_kernel<float>, grid, dim3(kBlock), 0, s,
                           (const ExchParams<float> *)b->exch[g][buf]);
    else
        hipLaunchKernelGGL(batch_exchange_kernel<double>, grid, dim3(kBlock), 0, s,
                           (const ExchParams<double> *)b->exch[g][buf]);
    TSMI_HIPB(hipGetLastError());
    return TSMI_OK;
}

// One step of group g's lattices from buffer cur[g]; the exchange when asked.
int single_pass(tsmi_batch *b, hipStream_t s, int g, int *cur, bool exch) {
    if (b->n_items[g] == 0) return TSMI_OK;
    const int st = b->esz == 4 ? single_typed<float, 4>(b, s, g, cur[g]) : single_typed<double, 2>(b, s, g, cur[g]);
    if (st) return st;
    cur[g] ^= 1;
    return exch ? exchange(b, s, g, cur[g]) : TSMI_OK;
}

// Two steps of group 1's lattices, then the 8-wide wrap.
int fused_pass(tsmi_batch *b, hipStream_t s, int *cur) {
    int st;
    if (b->esz == 4) st = b->lhy ? fused_mode<float, true>(b, s, cur[1]) : fused_mode<float, false>(b, s, cur[1]);
    else st = b->lhy ? fused_mode<double, true>(b, s, cur[1]) : fused_mode<double, false>(b, s, cur[1]);
    if (st) return st;
    cur[1] ^= 1;
    return exchange(b, s, 1, cur[1]);
}

// n steps of every lattice, eagerly; cur[] is updated to the buffers holding
// the results.  Group 0 goes to the side stream when both groups have work.
int batch_run(tsmi_batch *b, int n, int *cur) {
    const bool both = b->n_items[0] > 0 && b->n_items[1] > 0;
    hipStream_t s0 = both ? b->side : b->stream;
    if (both) {
        TSMI_HIPB(hipEventRecord(b->fork, b->stream));
        TSMI_HIPB(hipStreamWaitEvent(b->side, b->fork, 0));
    }
    int st = TSMI_OK;
    for (int i = 0; i < n && !st; ++i) st = single_pass(b, s0, 0, cur, i != n - 1);
    if (!st && b->n_items[1] > 0) {
        if (n < 3) {
            for (int i = 0; i < n && !st; ++i) st = single_pass(b, b->stream, 1, cur, i != n - 1);
        } else {
            st = single_pass(b, b->stream, 1, cur, true);
            int m = n - 2;
            for (; m >= 2 && !st; m -= 2) st = fused_pass(b, b->stream, cur);
            if (!st && m == 1) st = single_pass(b, b->stream, 1, cur, true);
            if (!st) st = single_pass(b, b->stream, 1, cur, false);
        }
    }
    if (both) {
        // join even after an error, so that a capture in progress stays well formed
        TSMI_HIPB(hipEventRecord(b->join, b->side));
        TSMI_HIPB(hipStreamWaitEvent(b->stream, b->join, 0));
    }
    return st;
}

// Resident blocks per CU of the persistent kernels, per E form (a persistent
// grid larger than what is resident would leave items to late waves).
template <typename T, bool IMAG, int VEC, bool LHY, bool PH>
hipError_t single_occupancy(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, batch_step_kernel<T, IMAG, VEC, LHY, 0, PH>, kBlock, 0);
}

template <bool PH>
hipError_t occupancy_of(const tsmi_batch *b, int *blocks) {
    if (b->esz == 4) {
        if (b->imag) return b->lhy ? single_occupancy<float, true, 4, true, PH>(blocks) : single_occupancy<float, true, 4, false, PH>(blocks);
        return b->lhy ? single_occupancy<float, false, 4, true, PH>(blocks) : single_occupancy<float, false, 4, false, PH>(blocks);
    }
    if (b->imag) return b->lhy ? single_occupancy<double, true, 2, true, PH>(blocks) : single_occupancy<double, true, 2, false, PH>(blocks);
    return b->lhy ? single_occupancy<double, false, 2, true, PH>(blocks) : single_occupancy<double, false, 2, false, PH>(blocks);
}

template <bool PH>
hipError_t occupancy2_of(const tsmi_batch *b, int *blocks) {
    auto q = [&](const void *k) { return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, kBlock, 0); };
    if (b->esz == 4)
        return b->lhy ? q((const void *)batch_step2_kernel<float, kVecF, true, 0, PH>)
                      : q((const void *)batch_step2_kernel<float, kVecF, false, 0, PH>);
    return b->lhy ? q((const void *)batch_step2_kernel<double, kVecF, true, 0, PH>)
                  : q((const void *)batch_step2_kernel<double, kVecF, false, 0, PH>);
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
    const int none[2] = {0, 0};
    TSMI_HIPB(hipMemcpyAsync(b->arena + L.part_n, none, sizeof(none), hipMemcpyHostToDevice, b->stream));
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

// The fused passes of a doubly periodic lattice read E over the whole 8-wide
// halo: the reference-tile table with every halo cell replaced by its wrapped
// inner cell.
int wrap_exp(tsmi_batch *b, const Lattice &L) {
    if (!L.halo8) return TSMI_OK;
    const size_t pb = (size_t)L.pitch * (size_t)L.Hd * (size_t)b->esz;
    for (int c = 0; c < 2; ++c)
        TSMI_HIPB(hipMemcpyAsync(b->arena + L.e8[c], b->arena + L.e[c] - L.shift, pb, hipMemcpyDeviceToDevice,
                                 b->stream));
    TSMI_HIPB(hipMemcpyAsync(b->arena + L.ph8, b->arena + L.ph - L.shift, pb, hipMemcpyDeviceToDevice, b->stream));
    const int hx = 2 * L.d.halo_x, hy = 2 * L.d.halo_y;
    const long cells = exchange_cells(L.Wd, L.Hd, 1, 1, hx, hy);
    const int eb = (int)std::min<long>(1024, (cells + 255) / 256);
    // E's two planes, then the one-plane form (given twice: the copy runs twice, to the same values)
    const size_t planes[2][2] = {{L.e8[0], L.e8[1]}, {L.ph8, L.ph8}};
    for (const auto &pl : planes) {
        if (b->esz == 4)
            hipLaunchKernelGGL((halo_exchange_kernel<float>), dim3(eb), dim3(256), 0, b->stream,
                               (float *)(b->arena + pl[0]), (float *)(b->arena + pl[1]), L.Wd, L.Hd, L.pitch, 1, 1,
                               hx, hy);
        else
            hipLaunchKernelGGL((halo_exchange_kernel<double>), dim3(eb), dim3(256), 0, b->stream,
                               (double *)(b->arena + pl[0]), (double *)(b->arena + pl[1]), L.Wd, L.Hd, L.pitch, 1,
                               1, hx, hy);
        TSMI_HIPB(hipGetLastError());
    }
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

}  // namespace

extern "C" {

int tsmi_batch_create(const tsmi_desc *descs, int32_t n, tsmi_batch **out) {
    if (!descs || !out || n < 1) return set_error(TSMI_E_INVALID, "tsmi_batch_create: need n >= 1 descriptors");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const tsmi_desc &d = descs[i];
        if (int st = check_desc(d, "tsmi_batch_create")) return st;
        if (d.precision != descs[0].precision || (d.imag_time != 0) != (descs[0].imag_time != 0) ||
            (d.coupling_lhy != 0.0) != (descs[0].coupling_lhy != 0.0))
            return set_error(TSMI_E_UNSUPPORTED,
                             "tsmi_batch_create: all lattices need the same precision, time direction and LHY setting");
        if (d.sweeps < 0 || d.sweeps > 2)
            return set_error(TSMI_E_INVALID, "tsmi_batch_create: sweeps must be 0, 1 or 2");
        if (d.sweeps == 2 && (d.imag_time || (d.periodic_x != 0) != (d.periodic_y != 0)))
            return set_error(TSMI_E_UNSUPPORTED, "tsmi_batch_create: two steps per pass need real time and both "
                                                 "axes closed or both periodic");
    }
    tsmi_batch *b = new tsmi_batch();
    b->n = n;
    b->esz = descs[0].precision == 32 ? 4 : 8;
    b->vec = b->esz == 4 ? 4 : 2;
    b->imag = descs[0].imag_time != 0;
    b->lhy = descs[0].coupling_lhy != 0.0;
    b->lat.resize(n);
    const int sw = kWave * b->vec - 2 * kApron;
    const int sw2 = kWave * kVecF - 4 * kApron;
    size_t off = 0;
    int mode = -2;
    for (int i = 0; i < n; ++i) {
        Lattice &L = b->lat[i];
        const tsmi_desc &d = descs[i];
        L.d = d;
        L.W = d.dim_x + (d.periodic_x ? 2 * d.halo_x : 0);
        L.H = d.dim_y + (d.periodic_y ? 2 * d.halo_y : 0);
        // 0 chooses two steps per pass wherever they are exact (the autotuner may
        // go back to one); 2 insists on them
        L.fusable = d.sweeps != 1 && !d.imag_time && (d.periodic_x != 0) == (d.periodic_y != 0);
        L.fused = L.fusable;
        L.tunable = L.fusable && d.sweeps == 0;
        L.halo8 = L.fusable && d.periodic_x;
        L.Wd = L.W + (L.halo8 ? 2 * d.halo_x : 0);
        L.Hd = L.H + (L.halo8 ? 2 * d.halo_y : 0);
        L.pitch = (L.Wd + 63) / 64 * 64;
        L.shift = L.halo8 ? ((size_t)d.halo_y * (size_t)L.pitch + (size_t)d.halo_x) * (size_t)b->esz : 0;
        L.n_strips = (L.W + sw - 1) / sw;
        L.fixed_rows = d.tile_rows > 0;
        set_rows(L, d.tile_rows > 0 ? d.tile_rows : kDefaultSegRows);
        if (L.fusable) {
            L.n_strips2 = (L.Wd + sw2 - 1) / sw2;
            set_rows2(L, d.tile_rows > 0 ? d.tile_rows : kDefaultSegRows);
        }
        L.max_waves = L.n_strips * ((L.H + 1) / 2);   // the smallest segment height, 2 rows
        L.ix0 = d.periodic_x ? d.halo_x : 0; L.ix1 = L.W - L.ix0;
        L.iy0 = d.periodic_y ? d.halo_y : 0; L.iy1 = L.H - L.iy0;
        L.k = step_coefs(d, L.H, b->esz);
        mode = mode == -2 ? L.k.mode : (mode == L.k.mode ? mode : -1);
        const size_t pb = align256((size_t)L.pitch * (size_t)L.Hd * (size_t)b->esz);
        for (int buf = 0; buf < 2; ++buf)
            for (int c = 0; c < 2; ++c) { L.psi[buf][c] = off + L.shift; off += pb; }
        for (int c = 0; c < 2; ++c) { L.e[c] = off + L.shift; off += pb; }
        for (int c = 0; c < 2; ++c) {
            if (L.halo8) { L.e8[c] = off; off += pb; }
            else L.e8[c] = L.e[c];
        }
        L.ph = off + L.shift; off += pb;
        if (L.halo8) { L.ph8 = off; off += pb; }
        else L.ph8 = L.ph;
        L.partials = off; off += align256(2 * (size_t)L.max_waves * sizeof(double));
        L.div = off; off += 256;
        L.part_n = off; off += 256;
    }
    b->mode = mode;
    b->arena_bytes = off;
    auto bail = [&](int code) {
        free_batch(b);
        delete b;
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("stream set-up: ") + hipGetErrorString(e)));
    b->stream = b->own;
    e = hipMalloc((void **)&b->arena, b->arena_bytes);
    if (e != hipSuccess) return bail(set_error(TSMI_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e)));
    if (e == hipSuccess) e = hipMemsetAsync(b->arena, 0, b->arena_bytes, b->stream);
    const size_t psz = b->esz == 4 ? sizeof(StepParams<float>) : sizeof(StepParams<double>);
    for (int p = 0; p < 2 && e == hipSuccess; ++p) {
        e = hipMalloc(&b->params[p], (size_t)n * psz);
        if (e == hipSuccess) e = hipMalloc(&b->params2[p], (size_t)n * psz);
    }
    if (e != hipSuccess) return bail(set_error(TSMI_E_OOM, std::string("batch allocation: ") + hipGetErrorString(e)));
    for (const Lattice &L : b->lat)
        if (int st = reset_divisor(b, L)) return bail(st);
    int st = b->esz == 4 ? upload_exchange<float>(b) : upload_exchange<double>(b);
    if (!st) st = upload_layout(b);
    if (st) return bail(st);
    int dev = 0, cus = 0, per_cu[2] = {0, 0}, per_cu2[2] = {0, 0};
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = occupancy_of<false>(b, &per_cu[0]);
    if (e == hipSuccess) e = occupancy_of<true>(b, &per_cu[1]);
    if (e == hipSuccess) e = occupancy2_of<false>(b, &per_cu2[0]);
    if (e == hipSuccess) e = occupancy2_of<true>(b, &per_cu2[1]);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e)));
    for (int f = 0; f < 2; ++f) {
        b->grid[f] = std::max(1, cus * std::max(1, per_cu[f]));
        b->grid2[f] = std::max(1, cus * std::max(1, per_cu2[f]));
    }
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bail(set_error(TSMI_E_HIP, std::string("batch set-up: ") + hipGetErrorString(e)));
    *out = b;
    return TSMI_OK;
}

void tsmi_batch_destroy(tsmi_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    free_batch(b);
    delete b;
}

int tsmi_batch_set_state(tsmi_batch *b, int32_t i, const double *re, const double *im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_set_state");
    if (!L) return TSMI_E_INVALID;
    if (!re || !im) return set_error(TSMI_E_INVALID, "tsmi_batch_set_state: null argument");
    const int buf = b->cur[L->group()];
    int st = upload_host_plane(b, *L, L->psi[buf][0], re);
    if (!st) st = upload_host_plane(b, *L, L->psi[buf][1], im);
    if (!st) st = reset_divisor(b, *L);
    return st;
}

int tsmi_batch_set_exp_potential(tsmi_batch *b, int32_t i, const double *e_re, const double *e_im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_set_exp_potential");
    if (!L) return TSMI_E_INVALID;
    if (!e_re || (!b->imag && !e_im)) return set_error(TSMI_E_INVALID, "tsmi_batch_set_exp_potential: null argument");
    // E times the frame correction, in FP64, rounded once
    const size_t n = (size_t)L->W * (size_t)L->H;
    std::vector<double> re(n), im(n, 0.0), ph(n, 0.0);
    bool ok = true;
    for (size_t k = 0; k < n; ++k) {
        const double er = e_re[k], ei = b->imag ? 0.0 : e_im[k];
        re[k] = er * L->k.corr_re - ei * L->k.corr_im;
        im[k] = er * L->k.corr_im + ei * L->k.corr_re;
        // the one-plane form, as the single-lattice import computes it
        if (!b->imag) {
            ok = ok && std::fabs(er * er + ei * ei - 1.0) <= 1e-12;
            ph[k] = -std::atan2(im[k], re[k]);
        } else {
            ok = ok && re[k] > 0.0;
            ph[k] = re[k] > 0.0 ? -std::log(re[k]) : 0.0;
        }
    }
    int st = upload_host_plane(b, *L, L->e[0], re.data());
    if (!st) st = upload_host_plane(b, *L, L->e[1], im.data());
    if (!st) st = upload_host_plane(b, *L, L->ph, ph.data());
    if (!st) st = wrap_exp(b, *L);
    if (st) return st;
    L->ph_ok = ok;
    bool all = true;
    for (const Lattice &M : b->lat) all = all && M.ph_ok;
    if (all != b->ph) {
        b->ph = all;
        drop_graphs(b);   // recorded with the other form's kernels
    }
    return TSMI_OK;
}

int tsmi_batch_evolve(tsmi_batch *b, int32_t n) {
    if (!b) return set_error(TSMI_E_INVALID, "tsmi_batch_evolve: null batch");
    if (n < 0) return set_error(TSMI_E_INVALID, "tsmi_batch_evolve: negative step count");
    if (n == 0) return TSMI_OK;
    const std::tuple<int, int, int> key(n, b->cur[0], b->cur[1]);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        GraphRec rec;
        rec.end[0] = b->cur[0];
        rec.end[1] = b->cur[1];
        TSMI_HIPB(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int st = batch_run(b, n, rec.end);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &captured);
        if (st) {
            if (captured) (void)hipGraphDestroy(captured);
            return st;
        }
        if (e != hipSuccess) return set_error(TSMI_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        rec.g = captured;
        const hipError_t ei = hipGraphInstantiate(&rec.x, rec.g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(rec.g);
            return set_error(TSMI_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        it = b->graphs.emplace(key, rec).first;
    }
    TSMI_HIPB(hipGraphLaunch(it->second.x, b->stream));
    b->cur[0] = it->second.end[0];
    b->cur[1] = it->second.end[1];
    return TSMI_OK;
}

int tsmi_batch_get_state(tsmi_batch *b, int32_t i, double *re, double *im) {
    Lattice *L = lattice_at(b, i, "tsmi_batch_get_state");
    if (!L) return TSMI_E_INVALID;
    if (!re || !im) return set_error(TSMI_E_INVALID, "tsmi_batch_get_state: null argument");
    const size_t n = (size_t)L->W * (size_t)L->H, es = (size_t)b->esz;
    const int buf = b->cur[L->group()];
    std::vector<unsigned char> tr(n * es), ti(n * es);
    double div = 1.0;
    TSMI_HIPB(hipMemcpy2DAsync(tr.data(), (size_t)L->W * es, b->arena + L->psi[buf][0], (size_t)L->pitch * es,
                               (size_t)L->W * es, L->H, hipMemcpyDeviceToHost, b->stream));
    TSMI_HIPB(hipMemcpy2DAsync(ti.data(), (size_t)L->W * es, b->arena + L->psi[buf][1], (size_t)L->pitch * es,
                               (size_t)L->W * es, L->H, hipMemcpyDeviceToHost, b->stream));
    if (b->imag) {
        hipLaunchKernelGGL(divisor_kernel<>, dim3(1), dim3(kWave), 0, b->stream,
                           (const double *)(b->arena + L->partials) + (size_t)buf * (size_t)L->max_waves,
                           (const int *)(b->arena + L->part_n) + buf, L->d.delta_x * L->d.delta_y, L->d.norm_target,
                           (double *)(b->arena + L->div));
        TSMI_HIPB(hipGetLastError());
        TSMI_HIPB(hipMemcpyAsync(&div, b->arena + L->div, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
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
    bool any_fusable = false, any_tunable = false;
    for (const Lattice &L : b->lat) {
        any_fusable = any_fusable || L.fusable;
        any_tunable = any_tunable || L.tunable;
    }
    // A trial reaches the fused passes only from 3 steps on, and a call of n steps
    // has 2 (or 3) single steps beside its fused passes: comparing one with two
    // steps per pass takes trials long enough for the passes to dominate.
    const int tsteps = any_tunable ? std::max(steps, 24) : (any_fusable ? std::max(steps, 6) : steps);
    // A lattice may change group: first bring both groups' states into the same
    // buffer, so that every membership finds its state where its group looks.
    if (b->cur[1] != b->cur[0]) {
        for (const Lattice &L : b->lat) {
            if (L.group() != 1) continue;
            const size_t pb = (size_t)L.pitch * (size_t)L.Hd * (size_t)b->esz;
            for (int c = 0; c < 2; ++c)
                TSMI_HIPB(hipMemcpyAsync(b->arena + L.psi[b->cur[0]][c] - L.shift, b->arena + L.psi[b->cur[1]][c] - L.shift,
                                         pb, hipMemcpyDeviceToDevice, b->stream));
        }
        b->cur[1] = b->cur[0];
        drop_graphs(b);
    }
    auto relayout = [&]() -> int {
        const int st = b->esz == 4 ? upload_exchange<float>(b) : upload_exchange<double>(b);
        return st ? st : upload_layout(b);
    };
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
    int cur[2] = {b->cur[0], b->cur[1]};
    auto trial = [&](float *ms) -> int {
        if (int st = relayout()) return st;
        if (int st = batch_run(b, 1, cur)) return st;   // warm-up
        TSMI_HIPB(hipEventRecord(t0, b->stream));
        if (int st = batch_run(b, tsteps, cur)) return st;
        TSMI_HIPB(hipEventRecord(t1, b->stream));
        TSMI_HIPB(hipEventSynchronize(t1));
        TSMI_HIPB(hipEventElapsedTime(ms, t0, t1));
        return TSMI_OK;
    };
    // size classes: lattices of the same kind (one step per pass, two, or
    // either) and tile shape, largest first
    std::map<std::tuple<int, int, int>, std::vector<int>> classes;
    for (int i = 0; i < b->n; ++i) {
        const Lattice &L = b->lat[i];
        if (!L.fixed_rows || L.tunable) classes[std::make_tuple(L.tunable ? 2 : L.group(), L.Hd, L.Wd)].push_back(i);
    }
    std::vector<std::tuple<int, int, int>> order;
    for (auto &c : classes) order.push_back(c.first);
    std::sort(order.begin(), order.end(), [](const std::tuple<int, int, int> &x, const std::tuple<int, int, int> &y) {
        return (long)std::get<1>(x) * std::get<2>(x) > (long)std::get<1>(y) * std::get<2>(y);
    });
    static const int cand[] = {8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
    int st = TSMI_OK;
    float best_all = 0.f;
    for (const auto &key : order) {
        const std::vector<int> &members = classes[key];
        const int kind = std::get<0>(key);
        const int height = std::get<1>(key);
        const Lattice &L0 = b->lat[members[0]];
        // the pass kinds to try: the class's own, or both for a tunable class;
        // the rows of the class's dominant pass, unless fixed by the descriptor
        std::vector<int> kinds = kind == 2 ? std::vector<int>{1, 0} : std::vector<int>{kind};
        auto apply = [&](int two, int rows) {
            for (int i : members) {
                Lattice &L = b->lat[i];
                L.fused = two == 1;
                if (L.fixed_rows) continue;
                if (two) set_rows2(L, rows);
                else set_rows(L, rows);
            }
        };
        // A trial times the whole batch, so a small class moves it little against
        // the timing noise: the class's current setting is timed first and a
        // candidate replaces the best only when it is faster by more than 1%.
        int best_kind = L0.fused ? 1 : 0;
        int best_rows = L0.fused ? L0.seg_rows2 : L0.seg_rows;
        float best = -1.f;
        apply(best_kind, best_rows);
        st = trial(&best);
        if (st) break;
        const int kind0 = best_kind, rows0 = best_rows;
        for (int two : kinds) {
            if (L0.fixed_rows && two == kind0) continue;   // already timed
            for (int rows : cand) {
                if (rows > height + 1 && rows != cand[0]) break;
                if (two == kind0 && rows == rows0) continue;
                apply(two, rows);
                float ms = 0.f;
                st = trial(&ms);
                if (st) break;
                if (ms < 0.99f * best) { best = ms; best_kind = two; best_rows = rows; }
                if (L0.fixed_rows) break;   // one trial per kind
            }
            if (st) break;
        }
        if (st) break;
        apply(best_kind, best_rows);
        best_all = best;
    }
    if (!st) st = relayout();
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    restore();
    if (st) return st;
    if (ms_per_step) *ms_per_step = best_all / tsteps;
    return TSMI_OK;
}

int tsmi_batch_info(const tsmi_batch *b, int32_t i, int32_t out[7]) {
    if (!b || !out || i < 0 || i >= b->n) return set_error(TSMI_E_INVALID, "tsmi_batch_info: no such lattice");
    const Lattice &L = b->lat[i];
    out[0] = L.W; out[1] = L.H; out[2] = L.pitch;
    if (L.fused) { out[3] = L.n_strips2; out[4] = L.n_segs2; out[5] = L.seg_rows2; }
    else { out[3] = L.n_strips; out[4] = L.n_segs; out[5] = L.seg_rows; }
    out[6] = L.fused ? 2 : 1;
    return TSMI_OK;
}

int tsmi_batch_synchronize(tsmi_batch *b) {
    if (!b) return set_error(TSMI_E_INVALID, "tsmi_batch_synchronize: null batch");
    TSMI_HIPB(hipStreamSynchronize(b->stream));
    return TSMI_OK;
}

}  // extern "C"
