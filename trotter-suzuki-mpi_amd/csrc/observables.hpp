// observables.hpp -- FP64 reductions over the tile: norm, energies and the
// State expected values, in one pass over HBM.
//
// The quantities are those of Solver::calculate_energy_expected_values
// (src/solver.cpp:320-420, single component, angular velocity 0) and of
// State::calculate_expected_values (src/model.cpp:408-455).
//
// Layout of the pass: a thread owns one column and walks a band of rows,
// keeping the five rows y-2..y+2 of its column in registers (the y stencils)
// and loading row y+3 one row ahead, so every state value is loaded once per
// band; the x stencils come from the neighbouring lanes through DPP lane
// shifts (a wave sums 60 columns, its two outer lanes on each side only supply
// stencil values).  Each thread sums 32 rows plainly and adds the chunk
// sums with Neumaier compensation; a wave combines its lanes in a fixed xor
// tree, a block its four waves in order, and obs_finalize_kernel sums the block
// partials with Neumaier compensation in block order -- the result does not
// depend on scheduling.  The sums are of the stored values; a pending
// imaginary-time divisor is applied to the final quantities (every sum is
// homogeneous in psi: degree 2, 4 for the interaction sum, 5 for LHY).
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernel.hpp"   // DPP lane shifts

namespace tsmi {

enum ObsIndex {
    OBS_N = 0,      // sum |psi|^2 over the inner region
    OBS_NK,         // sum |psi|^2 over the stencil margin
    OBS_LAPX,       // sum Re(conj(psi) D_x psi) over the margin, D = 4th-order 5-point
    OBS_LAPY,       // same along y
    OBS_POT,        // sum V |psi|^2
    OBS_INT,        // sum |psi|^4
    OBS_LHY,        // sum |psi|^5
    OBS_X, OBS_XX, OBS_Y, OBS_YY,
    OBS_PX,         // sum Im(conj(psi) d_x psi), one-sided 3rd-order derivative
    OBS_PY,
    OBS_LZ,
    OBS_COUNT
};

// what a pass sums: the norm alone, the energy sums (N .. LHY), or everything
enum ObsKind { OBS_KIND_NORM = 0, OBS_KIND_ENERGY = 1, OBS_KIND_ALL = 2 };

constexpr int kObsThreads = 256;
constexpr int kObsWaveCols = 60;         // columns summed per wave (64 lanes, 2 stencil lanes each side)
constexpr int kObsBlockCols = kObsThreads / 64 * kObsWaveCols;
constexpr int kObsChunk = 32;            // rows summed plainly before a compensated add
// Block partials a pass may write: the grid is (column blocks) x (row bands)
// with at most this many blocks, so the inner region may be at most
// kObsBlocksMax * kObsBlockCols columns wide.
constexpr int kObsBlocksMax = 4096;
constexpr int kObsOut = 16;              // obs_finalize_kernel's outputs

template <typename T>
struct ObsParams {
    const T *re, *im;
    const double *v;       // raw potential, same pitch (may be null: potential sums are 0)
    const double *xc, *yc; // mapped coordinates of tile columns / rows (OBS_KIND_ALL only)
    int W, H, pitch;
    int x0, x1, y0, y1;    // inner region
    int kx0, kx1, ky0, ky1;// stencil margin
    int rows_free;         // 1: the margin ignores the row bounds (energy with H == 1)
    int band;              // rows per block row
    double dx, dy;
    double *partials;      // [gridDim.y][gridDim.x][OBS_COUNT]
};

__device__ __forceinline__ void nsum(double &s, double &c, double v) {
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
}

template <int KIND>
constexpr int obs_used() {
    return KIND == OBS_KIND_NORM ? 1 : KIND == OBS_KIND_ENERGY ? OBS_LHY + 1 : OBS_COUNT;
}

template <typename T, int KIND, bool LHY>
__global__ void __launch_bounds__(kObsThreads) observables_kernel(ObsParams<T> p) {
    constexpr int NQ = obs_used<KIND>();
    double acc[NQ], s[NQ], c[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = s[i] = c[i] = 0.0;
    // lane l of a wave holds column x = base + l - 2; lanes 0, 1, 62, 63 only
    // supply the x stencils of their neighbours (DPP lane shifts)
    const int lane = threadIdx.x & 63;
    const int wave = (int)blockIdx.x * (kObsThreads / 64) + ((int)threadIdx.x >> 6);
    const int x = p.x0 + wave * kObsWaveCols + lane - 2;
    const bool col = lane >= 2 && lane < 62 && x < p.x1;
    const int xs = min(max(x, 0), p.W - 1);        // every lane loads a column of the tile
    const int yb0 = p.y0 + (int)blockIdx.y * p.band;
    const int yb1 = min(p.y1, yb0 + p.band);
    const long P = p.pitch;
    const bool col_margin = col && x >= p.kx0 && x < p.kx1;
    const double c1 = -1. / 12., c2 = 4. / 3., c3 = -2.5;
    const double d1 = 1. / 6., d2 = -1., d3 = 0.5, d4 = 1. / 3.;
    const double X = (KIND == OBS_KIND_ALL) ? p.xc[xs] : 0.0;
    // rows outside the tile read as row 0 or H-1: such values only meet cells
    // outside the stencil margin, whose sums do not use them
    auto row_of = [&](int y) { return (long)min(max(y, 0), p.H - 1) * P; };
    // rows y-2 .. y+2 of this column, and row y+3 loaded one row ahead
    double wr[5], wi[5], nr = 0.0, ni = 0.0, vn = 0.0;
    if (KIND == OBS_KIND_NORM) {
        const long o = row_of(yb0) + xs;
        nr = (double)p.re[o];
        ni = (double)p.im[o];
    } else {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const long o = row_of(yb0 - 2 + i) + xs;
            wr[i] = (double)p.re[o];
            wi[i] = (double)p.im[o];
        }
        if (p.v) vn = p.v[row_of(yb0) + xs];
    }
    int chunk = 0;
    for (int y = yb0; y < yb1; ++y) {
        double pr, pi, vk = 0.0;
        if (KIND == OBS_KIND_NORM) {
            pr = nr;
            pi = ni;
            const long o = row_of(y + 1) + xs;
            nr = (double)p.re[o];
            ni = (double)p.im[o];
        } else {
            const long o = row_of(y + 3) + xs;
            nr = (double)p.re[o];
            ni = (double)p.im[o];
            vk = vn;
            if (p.v) vn = p.v[row_of(y + 1) + xs];
            pr = wr[2];
            pi = wi[2];
        }
        const double n2 = col ? pr * pr + pi * pi : 0.0;
        acc[OBS_N] += n2;
        if constexpr (KIND != OBS_KIND_NORM) {
            acc[OBS_POT] += n2 * vk;
            acc[OBS_INT] += n2 * n2;
            if (LHY) acc[OBS_LHY] += n2 * n2 * sqrt(n2);
            if constexpr (KIND == OBS_KIND_ALL) {
                const double Y = p.yc[y];
                acc[OBS_X] += n2 * X;
                acc[OBS_XX] += n2 * X * X;
                acc[OBS_Y] += n2 * Y;
                acc[OBS_YY] += n2 * Y * Y;
            }
            // the x stencil from the neighbouring lanes (wave-uniform shifts)
            const double r1 = shfl_from_next(pr), i1 = shfl_from_next(pi);
            const double rr = shfl_from_next(r1), ir = shfl_from_next(i1);
            const double l1r = shfl_from_prev(pr), l1i = shfl_from_prev(pi);
            const double l2r = shfl_from_prev(l1r), l2i = shfl_from_prev(l1i);
            const bool row_margin = p.rows_free || (y >= p.ky0 && y < p.ky1);
            if (col_margin && row_margin) {
                acc[OBS_NK] += n2;
                const double lxr = c1 * rr + c2 * r1 + c2 * l1r + c1 * l2r + c3 * pr;
                const double lxi = c1 * ir + c2 * i1 + c2 * l1i + c1 * l2i + c3 * pi;
                acc[OBS_LAPX] += pr * lxr + pi * lxi;
                double dxr = 0.0, dxi = 0.0;
                if constexpr (KIND == OBS_KIND_ALL) {
                    dxr = d4 * r1 + d3 * pr + d2 * l1r + d1 * l2r;
                    dxi = d4 * i1 + d3 * pi + d2 * l1i + d1 * l2i;
                    acc[OBS_PX] += pr * dxi - pi * dxr;
                }
                if (p.H > 1) {
                    const double dr = wr[3], di = wi[3], ddr = wr[4], ddi = wi[4];
                    const double ur = wr[1], ui = wi[1], uur = wr[0], uui = wi[0];
                    const double lyr = c1 * ddr + c2 * dr + c2 * ur + c1 * uur + c3 * pr;
                    const double lyi = c1 * ddi + c2 * di + c2 * ui + c1 * uui + c3 * pi;
                    acc[OBS_LAPY] += pr * lyr + pi * lyi;
                    if constexpr (KIND == OBS_KIND_ALL) {
                        // "up" is row y-1 (src/model.cpp:448)
                        const double Y = p.yc[y];
                        const double dyr = d4 * ur + d3 * pr + d2 * dr + d1 * ddr;
                        const double dyi = d4 * ui + d3 * pi + d2 * di + d1 * ddi;
                        acc[OBS_PY] += pr * dyi - pi * dyr;
                        const double zr = Y / p.dx * dxr + X / p.dy * dyr;
                        const double zi = Y / p.dx * dxi + X / p.dy * dyi;
                        acc[OBS_LZ] += pr * zi - pi * zr;
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                wr[i] = wr[i + 1];
                wi[i] = wi[i + 1];
            }
            wr[4] = nr;
            wi[4] = ni;
        }
        if (++chunk == kObsChunk) {
            chunk = 0;
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                nsum(s[i], c[i], acc[i]);
                acc[i] = 0.0;
            }
        }
    }
    // the thread's total, the wave's in a fixed xor tree, the block's in wave order
    __shared__ double red[kObsThreads / 64][NQ];
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        nsum(s[i], c[i], acc[i]);
        double t = s[i] + c[i];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) red[wv][i] = t;
    }
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        double a = 0.0, cc = 0.0;
        for (int w = 0; w < kObsThreads / 64; ++w) nsum(a, cc, red[w][threadIdx.x]);
        const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
        p.partials[b * OBS_COUNT + threadIdx.x] = a + cc;
    }
}

// What obs_finalize_kernel turns the sums into (src/solver.cpp:538-561 for the
// energy, src/model.cpp:396-455 for the expected values).
struct ObsFinal {
    int kind;              // ObsKind
    int nq;                // quantities summed
    int blocks;            // block partials to add
    int rows2d;            // H > 1
    const double *div;     // pending imaginary-time divisor of the stored state (device; null: 1)
    double mass, g, g_lhy, dx, dy, length_x;
    int dim_x;
};

// One block of 64 * OBS_COUNT threads: wave i sums quantity i over the block
// partials (lane-strided Neumaier in block order, then a fixed xor tree), and
// thread 0 forms the outputs: OBS_KIND_NORM: out[0] = squared norm;
// OBS_KIND_ENERGY: out[0..5] as tsmi_energy; OBS_KIND_ALL: out[0..5] as
// tsmi_energy, out[6..15] as tsmi_expected_values.
__global__ void __launch_bounds__(64 * OBS_COUNT) obs_finalize_kernel(const double *partials, ObsFinal f,
                                                                      double *out) {
    __shared__ double q[OBS_COUNT];
    const int lane = threadIdx.x & 63, i = threadIdx.x >> 6;
    if (i < f.nq) {
        // sixteen loads per lane in flight before the first add; lanes past the
        // end add zeros, which leave their sum and compensation as they are
        constexpr int kBatch = 16;
        double a = 0.0, cc = 0.0;
        for (int b0 = 0; b0 < f.blocks; b0 += kBatch * 64) {
            double v[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const int b = b0 + j * 64 + lane;
                v[j] = b < f.blocks ? partials[(long)b * OBS_COUNT + i] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < kBatch; ++j) nsum(a, cc, v[j]);
        }
        double t = a + cc;
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) q[i] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double dv = f.div ? *f.div : 1.0;
    const double d2 = 1.0 / (dv * dv);
    const double n = q[OBS_N] * d2;
    if (f.kind == OBS_KIND_NORM) {
        out[0] = n * f.dx * f.dy;
        return;
    }
    // kinetic and potential terms are ratios of degree-2 sums: the divisor cancels
    const double cost = -1. / (2. * f.mass);
    const double kin = cost * (q[OBS_LAPX] / (f.dx * f.dx) + (f.rows2d ? q[OBS_LAPY] / (f.dy * f.dy) : 0.0)) /
                       q[OBS_NK];
    const double pot = q[OBS_POT] / q[OBS_N];
    const double itr = 0.5 * f.g * (q[OBS_INT] * d2) / q[OBS_N];
    const double lhy = 0.4 * f.g_lhy * (q[OBS_LHY] * d2 / dv) / q[OBS_N];
    out[0] = kin + pot + itr + lhy;                                   // src/solver.cpp:538
    out[1] = kin; out[2] = pot; out[3] = itr; out[4] = lhy;
    out[5] = n * f.dy * f.length_x / (double)f.dim_x;                 // src/solver.cpp:561
    if (f.kind != OBS_KIND_ALL) return;
    const double qn = q[OBS_N];
    const double ppx = -1. / f.dx, ppy = 1. / f.dy;                   // src/model.cpp:398-399
    out[6] = n * (f.dx * f.dy);
    out[7] = q[OBS_X] / qn; out[8] = q[OBS_XX] / qn;
    out[9] = q[OBS_Y] / qn; out[10] = q[OBS_YY] / qn;
    out[11] = -q[OBS_PX] * ppx / qn;
    out[12] = -q[OBS_LAPX] * ppx * ppx / qn;
    out[13] = -q[OBS_PY] * ppy / qn;
    out[14] = -q[OBS_LAPY] * ppy * ppy / qn;
    out[15] = q[OBS_LZ] / qn;
}

}  // namespace tsmi
