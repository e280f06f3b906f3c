// step_kernel.hpp -- the fused second-order Trotter-Suzuki step for gfx950.
//
// One wavefront owns a vertical strip of the tile: 64 lanes x VEC consecutive
// columns, of which the outer APRON=4 columns on each side are recomputed
// context (a step's horizontal dependency cone is 4 cells: 4 horizontal
// nearest-neighbour passes).  The wave streams down a segment of rows; the nine
// sub-passes V0 H0 V1 H1 P H1 V1 H0 V0 (src/cpukernel.cpp:22-56) run as a
// register pipeline in which each vertical pass is a stage that holds one
// pending row and emits the row above it once its lower partner has arrived, so
// a row leaves the pipeline 4 rows after it entered and nothing goes through LDS.
// Horizontal pairs inside a lane are register operations; the pair that straddles
// two lanes uses one cross-lane shuffle per value.  Each row is read once from
// HBM (plus the E row two rows behind it) and written once.
//
// Numerics follow the reference arithmetic (src/cpucartesian.cpp:4-138): the
// pair update [[a, ib], [ib, a]] (real time) or [[a, b], [b, a]] (imaginary time)
// on pairs (x,y)-(x+1,y) or (x,y)-(x,y+1) with x+y+s even in tile coordinates, and
// the potential step psi <- e^{-i(c n + c_lhy n^1.5)} E psi with n = |psi|^2 taken
// before the multiplication by E (imaginary time: psi <- exp(-(c n + c_lhy n^1.5)) Re(E) psi).
//
// The tile is the reference tile W x H (halo included on periodic axes).  The
// sweep is closed at the tile edges; for a periodic axis the halo cells carry the
// wrapped copies written by the previous step.  In "exchange" mode a step writes
// those copies itself (the single-process halo exchange, src/cpukernel.cpp:707-715
// then :738-746, including its row range H - 2*halo_y for the x halo); in the last
// step of an evolve call it writes the computed halo cells instead, as the
// reference does when it skips the exchange (src/solver.cpp:203-209).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace tsmi {

constexpr int kWave = 64;
constexpr int kApron = 4;

template <typename T, int VEC>
struct Row {
    T re[VEC];
    T im[VEC];
};

template <typename T>
struct StepParams {
    const T *in_re, *in_im;
    T *out_re, *out_im;
    const T *e_re, *e_im;        // E = exp(-i dt V) planes (e_im unused in imaginary time)
    int W, H, pitch;             // tile and row pitch (elements)
    int px, py, hx, hy;          // periodic flags and lattice halo widths
    int exchange;                // 1: write periodic halo copies, 0: write computed halo cells
    int n_strips, n_waves, seg_rows;
    T aH, bH, aV, bV, c_nl, c_lhy;
    int use_lhy;
    // imaginary time only
    const double *div_in;        // divisor of the previous renormalisation (nullptr: 1)
    double *div_out;             // divisor computed from this step's output
    double *partials;            // one FP64 partial sum per wave
    unsigned *counter;           // zero between launches
    double norm_target, dxdy;
    int ix0, ix1, iy0, iy1;      // inner region (tile coordinates)
};

// ---------------------------------------------------------------------------
// pair updates
// ---------------------------------------------------------------------------
template <bool IMAG, typename T>
__device__ __forceinline__ void rot(T a, T b, T &pr, T &pi, T &qr, T &qi) {
    const T r0 = pr, i0 = pi;
    if (!IMAG) {
        pr = a * r0 - b * qi;
        pi = a * i0 + b * qr;
        qr = a * qr - b * i0;
        qi = a * qi + b * r0;
    } else {
        pr = a * r0 + b * qr;
        pi = a * i0 + b * qi;
        qr = a * qr + b * r0;
        qi = a * qi + b * i0;
    }
}

// one side of a pair whose partner lives in another lane: update p from (q)
template <bool IMAG, typename T>
__device__ __forceinline__ void rot_p(T a, T b, T &pr, T &pi, T qr, T qi) {
    const T r0 = pr, i0 = pi;
    if (!IMAG) { pr = a * r0 - b * qi; pi = a * i0 + b * qr; }
    else       { pr = a * r0 + b * qr; pi = a * i0 + b * qi; }
}

// update q from (p)
template <bool IMAG, typename T>
__device__ __forceinline__ void rot_q(T a, T b, T pr, T pi, T &qr, T &qi) {
    if (!IMAG) { qr = a * qr - b * pi; qi = a * qi + b * pr; }
    else       { qr = a * qr + b * pr; qi = a * qi + b * pi; }
}

template <typename T>
__device__ __forceinline__ T shfl_from_next(T v) { return __shfl_down(v, 1, kWave); }
template <typename T>
__device__ __forceinline__ T shfl_from_prev(T v) { return __shfl_up(v, 1, kWave); }

// Horizontal coefficients of this lane.  A pair that is not inside the tile (or
// straddles the wave's edge) gets (1, 0), which leaves both cells unchanged.
template <typename T, int VEC>
struct HCoef {
    T ae[VEC / 2], be[VEC / 2];   // pairs (k, k+1), k even
    T ao[VEC / 2], bo[VEC / 2];   // pairs (k, k+1), k odd; index VEC/2-1 is (VEC-1, next lane's 0)
    T al, bl;                     // pair (previous lane's VEC-1, this lane's 0)
};

template <typename T, int VEC>
__device__ __forceinline__ HCoef<T, VEC> make_hcoef(int xl, int lane, int W, T a, T b) {
    HCoef<T, VEC> c;
    auto ok = [&](int x) { return x >= 0 && x + 1 < W; };
#pragma unroll
    for (int k = 0; k < VEC; k += 2) {
        const bool v = ok(xl + k);
        c.ae[k / 2] = v ? a : T(1);
        c.be[k / 2] = v ? b : T(0);
    }
#pragma unroll
    for (int k = 1; k < VEC; k += 2) {
        const bool v = ok(xl + k) && (k + 1 < VEC || lane < kWave - 1);
        c.ao[k / 2] = v ? a : T(1);
        c.bo[k / 2] = v ? b : T(0);
    }
    const bool vl = ok(xl - 1) && lane > 0;
    c.al = vl ? a : T(1);
    c.bl = vl ? b : T(0);
    return c;
}

// H(s) on row y: pairs (x, x+1) with x+y+s even; x has the parity of k, so the
// pattern is fixed by PAR = (y+s) mod 2, which the caller knows at compile time.
template <int PAR, bool IMAG, typename T, int VEC>
__device__ __forceinline__ void hpass(Row<T, VEC> &r, const HCoef<T, VEC> &c) {
    if constexpr (PAR == 0) {
#pragma unroll
        for (int k = 0; k < VEC; k += 2)
            rot<IMAG>(c.ae[k / 2], c.be[k / 2], r.re[k], r.im[k], r.re[k + 1], r.im[k + 1]);
    } else {
        const T nr = shfl_from_next(r.re[0]), ni = shfl_from_next(r.im[0]);
        const T pr = shfl_from_prev(r.re[VEC - 1]), pi = shfl_from_prev(r.im[VEC - 1]);
#pragma unroll
        for (int k = 1; k + 1 < VEC; k += 2)
            rot<IMAG>(c.ao[k / 2], c.bo[k / 2], r.re[k], r.im[k], r.re[k + 1], r.im[k + 1]);
        rot_p<IMAG>(c.ao[VEC / 2 - 1], c.bo[VEC / 2 - 1], r.re[VEC - 1], r.im[VEC - 1], nr, ni);
        rot_q<IMAG>(c.al, c.bl, pr, pi, r.re[0], r.im[0]);
    }
}

// V(s) between rows yu (upper) and yu+1 (lower): pairs with x+yu+s even,
// PAR = (yu+s) mod 2.
template <int PAR, bool IMAG, typename T, int VEC>
__device__ __forceinline__ void vpair(Row<T, VEC> &u, Row<T, VEC> &l, T a, T b) {
    if constexpr (PAR == 0) {
#pragma unroll
        for (int k = 0; k < VEC; k += 2) rot<IMAG>(a, b, u.re[k], u.im[k], l.re[k], l.im[k]);
    } else {
#pragma unroll
        for (int k = 1; k < VEC; k += 2) rot<IMAG>(a, b, u.re[k], u.im[k], l.re[k], l.im[k]);
    }
}

__device__ __forceinline__ void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ float exp_t(float x) { return expf(x); }
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }

// P on one row (src/cpucartesian.cpp:93-109 / :127-137).
template <bool IMAG, typename T, int VEC>
__device__ __forceinline__ void ppass(Row<T, VEC> &r, const Row<T, VEC> &e, T c, T cl, bool use_lhy) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T pr = r.re[k], pi = r.im[k];
        const T n2 = pr * pr + pi * pi;
        T ang = c * n2;
        if (use_lhy) ang = c * n2 + cl * (n2 * sqrt_t(n2));
        if (!IMAG) {
            T sn, cs;
            sincos_t(ang, &sn, &cs);
            const T r1 = e.re[k] * pr - e.im[k] * pi;
            const T i1 = e.re[k] * pi + e.im[k] * pr;
            r.re[k] = cs * r1 + sn * i1;
            r.im[k] = cs * i1 - sn * r1;
        } else {
            const T f = exp_t(T(-1) * ang);
            r.re[k] = f * e.re[k] * pr;
            r.im[k] = f * e.re[k] * pi;
        }
    }
}

// ---------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------
template <typename T, int VEC>
struct VecT;
template <> struct VecT<float, 4> { typedef float __attribute__((ext_vector_type(4))) type; };
template <> struct VecT<float, 2> { typedef float __attribute__((ext_vector_type(2))) type; };
template <> struct VecT<float, 8> { typedef float __attribute__((ext_vector_type(8))) type; };
template <> struct VecT<double, 2> { typedef double __attribute__((ext_vector_type(2))) type; };
template <> struct VecT<double, 4> { typedef double __attribute__((ext_vector_type(4))) type; };

template <typename T, int VEC>
__device__ __forceinline__ void load_vec(const T *p, T (&v)[VEC]) {
    typedef typename VecT<T, VEC>::type V;
    const V x = *reinterpret_cast<const V *>(p);
#pragma unroll
    for (int k = 0; k < VEC; ++k) v[k] = x[k];
}

template <typename T, int VEC>
__device__ __forceinline__ void store_vec(T *p, const T (&v)[VEC]) {
    typedef typename VecT<T, VEC>::type V;
    V x;
#pragma unroll
    for (int k = 0; k < VEC; ++k) x[k] = v[k];
    *reinterpret_cast<V *>(p) = x;
}

// ---------------------------------------------------------------------------
// halo bookkeeping for the exchange store rule
// ---------------------------------------------------------------------------
struct HaloMap {
    int W, H, px, py, hx, hy, nx, ny;   // nx = W - 2hx on a periodic x axis
    int rx0, rx1;                       // rows whose x halo the exchange refreshes
    __device__ bool xdest(int x, int y) const { return px && (x < hx || x >= W - hx) && y >= rx0 && y < rx1; }
    __device__ bool ydest(int y) const { return py && (y < hy || y >= H - hy); }
    // column receiving the x copy of source column x (or -1)
    __device__ int ximg(int x) const {
        if (!px) return -1;
        if (x >= hx && x < 2 * hx) return x + nx;
        if (x >= W - 2 * hx && x < W - hx) return x - nx;
        return -1;
    }
    __device__ int yimg(int y) const {
        if (!py) return -1;
        if (y >= hy && y < 2 * hy) return y + ny;
        if (y >= H - 2 * hy && y < H - hy) return y - ny;
        return -1;
    }
};

template <typename T>
__device__ __forceinline__ void put(T *re, T *im, long off, T vr, T vi) {
    re[off] = vr;
    im[off] = vi;
}

// Writes the final value of cell (x, y) and its periodic copies (exchange mode).
template <typename T>
__device__ __forceinline__ void store_cell_exchange(const StepParams<T> &p, const HaloMap &m, int x, int y, T vr, T vi) {
    const long row = (long)y * p.pitch;
    if (!m.xdest(x, y) && !m.ydest(y)) put(p.out_re, p.out_im, row + x, vr, vi);
    const bool xsrc_row = y >= m.rx0 && y < m.rx1;
    const int xi = xsrc_row ? m.ximg(x) : -1;
    if (xi >= 0 && !m.ydest(y)) put(p.out_re, p.out_im, row + xi, vr, vi);
    const int yi = m.yimg(y);
    if (yi >= 0) {
        const long yrow = (long)yi * p.pitch;
        if (!m.xdest(x, y)) put(p.out_re, p.out_im, yrow + x, vr, vi);
        if (xi >= 0) put(p.out_re, p.out_im, yrow + xi, vr, vi);
    }
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------
template <typename T, bool IMAG, int VEC>
__global__ void __launch_bounds__(256) step_kernel(StepParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wave >= p.n_waves) return;
    constexpr int SW = kWave * VEC - 2 * kApron;       // output columns per strip
    const int strip = wave % p.n_strips;
    const int seg = wave / p.n_strips;
    const int X0 = strip * SW - kApron;
    const int xl = X0 + lane * VEC;                   // first column of this lane
    const int y_out_lo = seg * p.seg_rows;
    const int y_out_hi = min(p.H, y_out_lo + p.seg_rows);
    const int y_lo = max(0, y_out_lo - kApron);
    const int y_hi = min(p.H, y_out_hi + kApron);

    // lanes that load: all VEC columns inside the row pitch
    const bool ld_ok = xl >= 0 && xl + VEC <= p.pitch;
    // lanes that store: not in the apron, inside the tile
    const bool own = lane >= kApron / VEC && lane < kWave - kApron / VEC;

    const HCoef<T, VEC> hc = make_hcoef<T, VEC>(xl, lane, p.W, p.aH, p.bH);
    const T aV = p.aV, bV = p.bV;

    double divd = 1.0;
    if (IMAG && p.div_in) divd = *p.div_in;

    HaloMap hm;
    hm.W = p.W; hm.H = p.H; hm.px = p.px; hm.py = p.py; hm.hx = p.hx; hm.hy = p.hy;
    hm.nx = p.W - 2 * p.hx; hm.ny = p.H - 2 * p.hy;
    hm.rx0 = p.py ? p.hy : 0;
    hm.rx1 = hm.rx0 + max(0, p.H - 2 * p.hy);
    const bool strip_plain = !p.exchange || !p.px ||
                             (X0 + kApron >= 2 * p.hx && X0 + kApron + SW <= p.W - 2 * p.hx);

    double acc = 0.0, comp = 0.0;   // Neumaier sum of |psi|^2 over stored inner cells

    auto load_psi = [&](int y, Row<T, VEC> &r) {
        if (ld_ok) {
            const long off = (long)y * p.pitch + xl;
            load_vec<T, VEC>(p.in_re + off, r.re);
            load_vec<T, VEC>(p.in_im + off, r.im);
            if (IMAG && p.div_in) {
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    r.re[k] = T(double(r.re[k]) / divd);
                    r.im[k] = T(double(r.im[k]) / divd);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) { r.re[k] = T(0); r.im[k] = T(0); }
        }
    };
    auto load_e = [&](int y, Row<T, VEC> &r) {
        if (ld_ok && y >= y_lo && y < y_hi) {
            const long off = (long)y * p.pitch + xl;
            load_vec<T, VEC>(p.e_re + off, r.re);
            if (!IMAG) load_vec<T, VEC>(p.e_im + off, r.im);
            else {
#pragma unroll
                for (int k = 0; k < VEC; ++k) r.im[k] = T(0);
            }
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) { r.re[k] = T(1); r.im[k] = T(0); }
        }
    };

    Row<T, VEC> pend1, pend2, pend3, pend4;
    Row<T, VEC> inA, inB, eA, eB;

    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };

    // One pipeline beat for input row t; TP = t mod 2 fixes every pair pattern
    // (rows handled by a wave start at an even row, see the host launcher).
    auto process = [&](auto tp, int t, Row<T, VEC> &in, const Row<T, VEC> &e) {
        constexpr int TP = decltype(tp)::value;
        // stage 1: V0, emits row t-1
        Row<T, VEC> o1;
        if (valid(t) && valid(t - 1)) vpair<(TP + 1) & 1, IMAG>(pend1, in, aV, bV);
        o1 = pend1;
        if (valid(t)) pend1 = in;
        if (valid(t - 1)) hpass<(TP + 1) & 1, IMAG>(o1, hc);
        // stage 2: V1, emits row t-2
        Row<T, VEC> o2;
        if (valid(t - 1) && valid(t - 2)) vpair<(TP + 2 + 1) & 1, IMAG>(pend2, o1, aV, bV);
        o2 = pend2;
        if (valid(t - 1)) pend2 = o1;
        if (valid(t - 2)) {
            hpass<(TP + 2 + 1) & 1, IMAG>(o2, hc);
            ppass<IMAG>(o2, e, p.c_nl, p.c_lhy, p.use_lhy != 0);
            hpass<(TP + 2 + 1) & 1, IMAG>(o2, hc);
        }
        // stage 3: V1, emits row t-3
        Row<T, VEC> o3;
        if (valid(t - 2) && valid(t - 3)) vpair<(TP + 3 + 1) & 1, IMAG>(pend3, o2, aV, bV);
        o3 = pend3;
        if (valid(t - 2)) pend3 = o2;
        if (valid(t - 3)) hpass<(TP + 3) & 1, IMAG>(o3, hc);
        // stage 4: V0, emits row t-4
        Row<T, VEC> o4;
        if (valid(t - 3) && valid(t - 4)) vpair<(TP + 4) & 1, IMAG>(pend4, o3, aV, bV);
        o4 = pend4;
        if (valid(t - 3)) pend4 = o3;
        const int y = t - 4;
        if (y >= y_out_lo && y < y_out_hi && own) {
            const bool row_plain = strip_plain &&
                (!p.exchange || !p.py || (y >= 2 * p.hy && y < p.H - 2 * p.hy));
            if (row_plain && xl + VEC <= p.W) {
                const long off = (long)y * p.pitch + xl;
                store_vec<T, VEC>(p.out_re + off, o4.re);
                store_vec<T, VEC>(p.out_im + off, o4.im);
            } else {
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const int x = xl + k;
                    if (x < 0 || x >= p.W) continue;
                    if (!p.exchange) put(p.out_re, p.out_im, (long)y * p.pitch + x, o4.re[k], o4.im[k]);
                    else store_cell_exchange(p, hm, x, y, o4.re[k], o4.im[k]);
                }
            }
            if (IMAG && y >= p.iy0 && y < p.iy1) {
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const int x = xl + k;
                    if (x >= p.ix0 && x < p.ix1) {
                        const double vr = double(o4.re[k]), vi = double(o4.im[k]);
                        const double v = vr * vr + vi * vi;
                        const double t2 = acc + v;
                        comp += (fabs(acc) >= fabs(v)) ? (acc - t2) + v : (v - t2) + acc;
                        acc = t2;
                    }
                }
            }
        }
    };

    // software pipeline: rows t+1 and t+2 (and E rows t-1, t) are in flight while row t runs
    int t = y_lo;
    if (t < y_hi) load_psi(t, inA);
    load_e(t - 2, eA);
    if (t + 1 < y_hi) load_psi(t + 1, inB);
    load_e(t - 1, eB);
    const int t_end = y_hi + 4;
    for (; t < t_end; t += 2) {
        process(std::integral_constant<int, 0>(), t, inA, eA);
        if (t + 2 < y_hi) load_psi(t + 2, inA);
        load_e(t, eA);
        if (t + 1 < t_end) {
            process(std::integral_constant<int, 1>(), t + 1, inB, eB);
            if (t + 3 < y_hi) load_psi(t + 3, inB);
            load_e(t + 1, eB);
        }
    }

    if (IMAG) {
        // deterministic renormalisation: one partial per wave; the wave that
        // finishes last sums them in wave order and publishes the divisor
        double s = acc + comp;
        for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
        unsigned ticket = 0u;
        if (lane == 0) {
            p.partials[wave] = s;
            __threadfence();
            ticket = atomicAdd(p.counter, 1u);
        }
        ticket = __shfl(ticket, 0, kWave);
        if (ticket == (unsigned)p.n_waves - 1) {
            __threadfence();
            double a2 = 0.0, c2 = 0.0;
            for (int i = lane; i < p.n_waves; i += kWave) {
                const double v = __hip_atomic_load(p.partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const double t2 = a2 + v;
                c2 += (fabs(a2) >= fabs(v)) ? (a2 - t2) + v : (v - t2) + a2;
                a2 = t2;
            }
            double tot = a2 + c2;
            for (int o = kWave / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
            if (lane == 0) {
                const double nrm = tot * p.dxdy;
                *p.div_out = (p.norm_target != 0.0) ? sqrt(nrm / p.norm_target) : 1.0;
                __hip_atomic_store(p.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

}  // namespace tsmi
