// batch_kernel.hpp -- many independent lattices advanced by one launch per step.
//
// Each state of a batch keeps its own tile, E planes and scaled-frame factors
// (one StepParams per state and buffer parity, in device memory).  A step is
// one persistent launch: the resident wavefronts walk the work items -- a
// (state, strip-segment) pair, the unit step_kernel gives one wave -- with a
// wave-uniform stride, so lattices of very different sizes share the launch
// without a partial last round.  The items are sorted by decreasing cost on the
// host, so the long ones are spread over all waves first.  The item index is
// scalar arithmetic from the wave's own index: the loop's trip count is uniform
// by construction, which keeps every lane of a wave on the same item through
// the divergent parts of the step (the imaginary-time reduction).
//
// batch_step2_kernel is the same walk over the work items of the fused
// decomposition (fused_kernel.hpp): two real-time steps per pass.
#pragma once
#include "fused_kernel.hpp"
#include "step_kernel.hpp"

namespace tsmi {

template <typename T, bool IMAG, int VEC, bool LHY, int MODE, bool PH>
__global__ void __launch_bounds__(256) batch_step_kernel(const StepParams<T> *__restrict__ params,
                                                         const int2 *__restrict__ items, int n_items) {
    const int lane = threadIdx.x & (kWave - 1);
    const int first = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave));
    const int stride = __builtin_amdgcn_readfirstlane(gridDim.x * (blockDim.x / kWave));
    for (int it = first; it < n_items; it += stride) {
        const int2 e = items[it];
        step_wave<T, IMAG, VEC, LHY, MODE, PH>(params[e.x], e.y, lane);
    }
}

template <typename T, int VEC, bool LHY, int MODE, bool PH>
__global__ void __launch_bounds__(256) batch_step2_kernel(const StepParams<T> *__restrict__ params,
                                                          const int2 *__restrict__ items, int n_items) {
    const int lane = threadIdx.x & (kWave - 1);
    const int first = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave));
    const int stride = __builtin_amdgcn_readfirstlane(gridDim.x * (blockDim.x / kWave));
    for (int it = first; it < n_items; it += stride) {
        const int2 e = items[it];
        step2_wave<T, VEC, LHY, MODE, PH>(params[e.x], e.y, lane);
    }
}

template <typename T>
struct ExchParams {
    T *re, *im;
    int W, H, pitch, px, py, hx, hy;
};

// gridDim.y workgroups per periodic state (blockIdx.x names the state): its
// halo cells in grid-stride order.
template <typename T>
__global__ void __launch_bounds__(256) batch_exchange_kernel(const ExchParams<T> *__restrict__ list) {
    const ExchParams<T> q = list[blockIdx.x];
    const int n = exchange_cells(q.W, q.H, q.px, q.py, q.hx, q.hy);
    for (int i = blockIdx.y * blockDim.x + threadIdx.x; i < n; i += gridDim.y * blockDim.x)
        exchange_cell<T>(i, q.re, q.im, q.W, q.H, q.pitch, q.px, q.py, q.hx, q.hy);
}

}  // namespace tsmi
