// observables.hpp -- FP64 reductions over the tile: norm, energies and the
// State expected values, in one pass.
//
// Each block sums a stripe of rows with per-thread Neumaier compensation, then
// the block's threads are summed in thread order, and the host sums the block
// partials in block order, so the result does not depend on scheduling.
// The quantities are those of Solver::calculate_energy_expected_values
// (src/solver.cpp:320-420, single component, angular velocity 0) and of
// State::calculate_expected_values (src/model.cpp:408-455).
#pragma once
#include <hip/hip_runtime.h>

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

constexpr int kObsThreads = 256;

template <typename T>
struct ObsParams {
    const T *re, *im;
    const double *v;       // raw potential, same pitch (may be null: potential sums are 0)
    const double *xc, *yc; // mapped coordinates of tile columns / rows
    int W, H, pitch;
    int x0, x1, y0, y1;    // inner region
    int kx0, kx1, ky0, ky1;// stencil margin
    int rows_free;         // 1: the margin ignores the row bounds (energy with H == 1)
    double div;            // loaded values are divided by this (pending renormalisation)
    double dx, dy;
    double *partials;      // [gridDim.x][OBS_COUNT]
};

__device__ __forceinline__ void nsum(double &s, double &c, double v) {
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
}

template <typename T>
__global__ void __launch_bounds__(kObsThreads) observables_kernel(ObsParams<T> p) {
    double s[OBS_COUNT], c[OBS_COUNT];
#pragma unroll
    for (int i = 0; i < OBS_COUNT; ++i) { s[i] = 0.0; c[i] = 0.0; }
    const double c1 = -1. / 12., c2 = 4. / 3., c3 = -2.5;
    const double d1 = 1. / 6., d2 = -1., d3 = 0.5, d4 = 1. / 3.;
    const bool has_v = p.v != nullptr;
    for (int y = p.y0 + (int)blockIdx.x; y < p.y1; y += (int)gridDim.x) {
        const long row = (long)y * p.pitch;
        const double Y = p.yc[y];
        const bool row_margin = p.rows_free || (y >= p.ky0 && y < p.ky1);
        for (int x = p.x0 + (int)threadIdx.x; x < p.x1; x += kObsThreads) {
            auto R = [&](long o) { return double(p.re[o]) / p.div; };
            auto I = [&](long o) { return double(p.im[o]) / p.div; };
            const long k = row + x;
            const double pr = R(k), pi = I(k);
            const double n2 = pr * pr + pi * pi;
            const double X = p.xc[x];
            nsum(s[OBS_N], c[OBS_N], n2);
            if (has_v) nsum(s[OBS_POT], c[OBS_POT], n2 * p.v[k]);
            nsum(s[OBS_INT], c[OBS_INT], n2 * n2);
            nsum(s[OBS_LHY], c[OBS_LHY], pow(n2, 2.5));
            nsum(s[OBS_X], c[OBS_X], n2 * X);
            nsum(s[OBS_XX], c[OBS_XX], n2 * X * X);
            nsum(s[OBS_Y], c[OBS_Y], n2 * Y);
            nsum(s[OBS_YY], c[OBS_YY], n2 * Y * Y);
            if (!(row_margin && x >= p.kx0 && x < p.kx1)) continue;
            nsum(s[OBS_NK], c[OBS_NK], n2);
            const long P = p.pitch;
            const double rr = R(k + 2), ir = I(k + 2), r1 = R(k + 1), i1 = I(k + 1);
            const double l1r = R(k - 1), l1i = I(k - 1), l2r = R(k - 2), l2i = I(k - 2);
            const double lxr = c1 * rr + c2 * r1 + c2 * l1r + c1 * l2r + c3 * pr;
            const double lxi = c1 * ir + c2 * i1 + c2 * l1i + c1 * l2i + c3 * pi;
            nsum(s[OBS_LAPX], c[OBS_LAPX], pr * lxr + pi * lxi);
            const double dxr = d4 * r1 + d3 * pr + d2 * l1r + d1 * l2r;
            const double dxi = d4 * i1 + d3 * pi + d2 * l1i + d1 * l2i;
            nsum(s[OBS_PX], c[OBS_PX], pr * dxi - pi * dxr);
            if (p.H > 1) {
                const double dr = R(k + P), di = I(k + P), ddr = R(k + 2 * P), ddi = I(k + 2 * P);
                const double ur = R(k - P), ui = I(k - P), uur = R(k - 2 * P), uui = I(k - 2 * P);
                const double lyr = c1 * ddr + c2 * dr + c2 * ur + c1 * uur + c3 * pr;
                const double lyi = c1 * ddi + c2 * di + c2 * ui + c1 * uui + c3 * pi;
                nsum(s[OBS_LAPY], c[OBS_LAPY], pr * lyr + pi * lyi);
                // "up" is row y-1 (src/model.cpp:448)
                const double dyr = d4 * ur + d3 * pr + d2 * dr + d1 * ddr;
                const double dyi = d4 * ui + d3 * pi + d2 * di + d1 * ddi;
                nsum(s[OBS_PY], c[OBS_PY], pr * dyi - pi * dyr);
                const double zr = Y / p.dx * dxr + X / p.dy * dyr;
                const double zi = Y / p.dx * dxi + X / p.dy * dyi;
                nsum(s[OBS_LZ], c[OBS_LZ], pr * zi - pi * zr);
            }
        }
    }
    __shared__ double red[OBS_COUNT][kObsThreads];
#pragma unroll
    for (int i = 0; i < OBS_COUNT; ++i) red[i][threadIdx.x] = s[i] + c[i];
    __syncthreads();
    if (threadIdx.x < OBS_COUNT) {
        double a = 0.0, cc = 0.0;
        for (int t = 0; t < kObsThreads; ++t) nsum(a, cc, red[threadIdx.x][t]);
        p.partials[(long)blockIdx.x * OBS_COUNT + threadIdx.x] = a + cc;
    }
}

}  // namespace tsmi
