// batch_kernel.hpp -- many independent lattices advanced by one launch per step.
//
// Each state of a batch keeps its own tile, E planes and scaled-frame factors
// (one StepParams per state and buffer parity, in device memory).  A step is
// one persistent launch: every resident wavefront takes work items -- a
// (state, strip-segment) pair, the unit step_kernel gives one wave -- from a
// queue counter until the queue is empty, so lattices of very different sizes
// share the launch without a partial last round.  The items are sorted by
// decreasing cost on the host, so the long ones start first.  Counters are used
// in alternate launches: a launch draws from queue[parity] and clears
// queue[1 - parity], which the previous launch used and the next one will, so
// no host-side reset (and no memset node in a captured graph) is needed.
#pragma once
#include "step_kernel.hpp"

namespace tsmi {

template <typename T, bool IMAG, int VEC, bool LHY, int MODE>
__global__ void __launch_bounds__(256) batch_step_kernel(const StepParams<T> *__restrict__ params,
                                                         const int2 *__restrict__ items, int n_items,
                                                         unsigned *__restrict__ queue, int parity) {
    const int lane = threadIdx.x & (kWave - 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) queue[1 - parity] = 0u;
    for (;;) {
        unsigned it = 0u;
        if (lane == 0) it = atomicAdd(queue + parity, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= (unsigned)n_items) break;   // every wave ends here once the queue is drained
        const int2 e = items[it];
        step_wave<T, IMAG, VEC, LHY, MODE>(params[e.x], e.y, lane);
    }
}

template <typename T>
struct ExchParams {
    T *re, *im;
    int W, H, pitch, px, py, hx, hy;
};

// One workgroup per periodic state: its halo cells in workgroup-stride order.
template <typename T>
__global__ void __launch_bounds__(256) batch_exchange_kernel(const ExchParams<T> *__restrict__ list) {
    const ExchParams<T> q = list[blockIdx.x];
    const int n = exchange_cells(q.W, q.H, q.px, q.py, q.hx, q.hy);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        exchange_cell<T>(i, q.re, q.im, q.W, q.H, q.pitch, q.px, q.py, q.hx, q.hy);
}

}  // namespace tsmi
