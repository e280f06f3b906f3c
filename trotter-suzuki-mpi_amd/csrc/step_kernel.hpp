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
// sweep is closed at the tile edges and writes every cell; on a periodic axis the
// halo cells then receive the wrapped copies from halo_exchange_kernel (the
// single-process halo exchange, src/cpukernel.cpp:707-715 then :738-746), which
// the host skips after the last step of an evolve call as the reference does
// (src/solver.cpp:203-209).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace tsmi {

constexpr int kWave = 64;
constexpr int kApron = 4;

// XCD-aware workgroup order.  The dispatcher hands workgroup b to XCD b mod 8
// (8 XCDs on gfx950, each with its own L2).  xcd_block<true>(b, n) gives every
// XCD a contiguous range of logical workgroups: the neighbouring strips of a
// segment -- which read the same apron columns and share the cache lines at
// their edges, at about the same time -- then sit behind one L2 instead of
// fetching those lines once per XCD.  A bijection of [0, n) for any n (and any
// XCD count).  Each kernel turns it on where it measured faster (8192^2: the
// FP32 two-step pass fetches 8.8% fewer bytes and runs 1% faster, the FP64 one
// 1.5% slower; the one-step pass 2.4% slower in FP32, 2.3% faster in FP64; the
// two-component passes 0.4-2.5% faster).
constexpr int kXcds = 8;
template <bool ON>
__device__ __forceinline__ int xcd_block(int b, int n) {
    if constexpr (!ON) {
        return b;
    } else {
        const int q = n / kXcds, r = n - q * kXcds;
        const int x = b - (b / kXcds) * kXcds;
        return x * q + min(x, r) + b / kXcds;
    }
}

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
    // The same factor as one plane, when every |E| = 1 (real time) or E > 0
    // (imaginary time): phi = -arg(E corr), or lambda = -ln(E corr); the step
    // then applies mag exp(-i (phase + phi)), or exp(-(decay + lambda)).
    const T *ph;
    double mag;                  // |corr| (real time, phase form)
    int W, H, pitch;             // tile and row pitch (elements)
    int n_strips, n_waves, seg_rows;
    // Pair factors in the scaled frame (see the note at fpair): a pair becomes
    // p' = m p + i t q (real time) or m p + t q (imaginary time), with (m, t) =
    // (1, b/a) in mode A and (a/b, 1) in mode B; a cell without a partner in a
    // pass is multiplied by un (= 1/g of that axis).
    T mH, tH, mV, tV, unH, unV;
    T c_nl, c_lhy;               // nonlinear couplings, already rescaled to the frame
    // Imaginary time only.  The renormalisation of a step's output is applied
    // as the next step loads it: every wave of a step forms its FP64 partial
    // sum of |psi|^2, stored per wave (batches) or combined per workgroup
    // (step_kernel), with their count, and every wave of the next launch sums
    // them in order (renorm_divisor).  The launch boundary makes the stores
    // visible; no wave waits for another workgroup inside a launch.
    const double *part_in;       // partial sums of the step that wrote the input (nullptr: divisor 1)
    const int *n_in;             // their count (0: divisor 1)
    double *part_out;            // this step's partial sums, one per wave or per workgroup
    int *n_out;                  // and their count
    double norm_target, dxdy;
    int ix0, ix1, iy0, iy1;      // inner region (tile coordinates)
};

// The imaginary-time divisor sqrt(N / N_target), N = dxdy * (sum of a step's
// per-wave partials), or 1 for no partials or no target (src/cpukernel.cpp:
// 548-560 divides by it).  Lane-strided Neumaier sums in wave order, then an
// xor butterfly: partner lanes add the same two operands, so every lane of the
// wave returns the same value, and every caller the same bits.  The partials
// are read eight per lane at a time (all loads issued before the first wait);
// a lane past the end adds zeros, which leave its sum and compensation as they are.
__device__ __forceinline__ double renorm_divisor(const double *part, int n, double dxdy, double target, int lane) {
    constexpr int kBatch = 8;
    double a = 0.0, c = 0.0;
    for (int i0 = 0; i0 < n; i0 += kBatch * kWave) {
        double x[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) x[j] = part[min(i0 + j * kWave + lane, n - 1)];
        __builtin_amdgcn_sched_barrier(0);   // every load of the batch before the first use
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const double v = i0 + j * kWave + lane < n ? x[j] : 0.0;
            const double t = a + v;
            c += (fabs(a) >= fabs(v)) ? (a - t) + v : (v - t) + a;
            a = t;
        }
    }
    double tot = a + c;
    for (int o = kWave / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
    if (n <= 0 || target == 0.0) return 1.0;
    const double nrm = tot * dxdy;
    return sqrt(nrm / target);
}

// Divides a row by the pending renormalisation divisor d: a multiplication by
// 1/d in the row's precision.  The reciprocal's rounding scales the whole state
// by one factor (within half an ulp of 1), which the next renormalisation
// removes; each element is rounded once, as a quotient would be.
template <typename T, int VEC>
__device__ __forceinline__ void scale_row(Row<T, VEC> &r, double inv) {
    const T s = T(inv);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        r.re[k] *= s;
        r.im[k] *= s;
    }
}

// One wave: the divisor of a pending renormalisation, for host-side readers
// (a template only so that every source including this header shares it).
template <typename = void>
__global__ void __launch_bounds__(64) divisor_kernel(const double *part, const int *n, double dxdy, double target,
                                                     double *out) {
    const int lane = threadIdx.x;
    const double d = renorm_divisor(part, __builtin_amdgcn_readfirstlane(*n), dxdy, target, lane);
    if (lane == 0) *out = d;
}

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

// Scaled frame.  The exact pair factor a I + i b sigma_x equals s (m I + i t sigma_x)
// with s = a (mode A, |a| >= |b|) or s = b (mode B); the kernel applies only the
// bracket, whose one coefficient is 1, so each output is a single fused
// multiply-add.  The dropped per-pass scales are uniform, so the host folds
// their product for a whole step into E and rescales the nonlinear coupling to
// the density seen mid-step; with the coefficients rounded to the kernel's
// precision it uses g = (exact eigenvalue)/(computed eigenvalue) of the smooth
// modes per pass, which keeps the modulus exact for every mode.
template <bool IMAG, int MODE, typename T>
__device__ __forceinline__ void fpair(T k, T &pr, T &pi, T &qr, T &qi) {
    const T r0 = pr, i0 = pi;
    if (!IMAG) {
        if (MODE == 0) {          // p + i k q
            pr = fma(-k, qi, r0); pi = fma(k, qr, i0);
            qr = fma(-k, i0, qr); qi = fma(k, r0, qi);
        } else {                  // k p + i q
            pr = fma(k, r0, -qi); pi = fma(k, i0, qr);
            qr = fma(k, qr, -i0); qi = fma(k, qi, r0);
        }
    } else {                      // p + k q (mode A only)
        pr = fma(k, qr, r0); pi = fma(k, qi, i0);
        qr = fma(k, r0, qr); qi = fma(k, i0, qi);
    }
}

template <bool IMAG, int MODE, typename T>
__device__ __forceinline__ void fpair_p(T k, T &pr, T &pi, T qr, T qi) {
    const T r0 = pr, i0 = pi;
    if (!IMAG) {
        if (MODE == 0) { pr = fma(-k, qi, r0); pi = fma(k, qr, i0); }
        else           { pr = fma(k, r0, -qi); pi = fma(k, i0, qr); }
    } else { pr = fma(k, qr, r0); pi = fma(k, qi, i0); }
}

template <bool IMAG, int MODE, typename T>
__device__ __forceinline__ void fpair_q(T k, T pr, T pi, T &qr, T &qi) {
    if (!IMAG) {
        if (MODE == 0) { qr = fma(-k, pi, qr); qi = fma(k, pr, qi); }
        else           { qr = fma(k, qr, -pi); qi = fma(k, qi, pr); }
    } else { qr = fma(k, pr, qr); qi = fma(k, pi, qi); }
}

// Neighbour-lane values through DPP whole-wave shifts by one lane (no LDS
// traffic): lane i gets lane i+1's value (from_next) or lane i-1's (from_prev);
// the lane without a neighbour gets 0 (bound control: no old value to set
// first, so the move can also fold into the instruction that reads it), which
// only ever reaches apron columns.
__device__ __forceinline__ int dpp_from_next(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_from_prev(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ float shfl_from_next(float v) {
    return __builtin_bit_cast(float, dpp_from_next(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float shfl_from_prev(float v) {
    return __builtin_bit_cast(float, dpp_from_prev(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double shfl_from_next(double v) {
    const int2 b = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(dpp_from_next(b.x), dpp_from_next(b.y)));
}
__device__ __forceinline__ double shfl_from_prev(double v) {
    const int2 b = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(dpp_from_prev(b.x), dpp_from_prev(b.y)));
}

// Horizontal coefficients of this lane.  A pair that is not inside the tile (or
// straddles the wave's edge) gets (1, 0), which leaves both cells unchanged.
template <typename T, int VEC>
struct HCoef {
    T ae[VEC / 2], be[VEC / 2];   // pairs (k, k+1), k even
    T ao[VEC / 2], bo[VEC / 2];   // pairs (k, k+1), k odd; index VEC/2-1 is (VEC-1, next lane's 0)
    T al, bl;                     // pair (previous lane's VEC-1, this lane's 0)
};

template <typename T, int VEC>
__device__ __forceinline__ HCoef<T, VEC> make_hcoef(int xl, int lane, int W, T a, T b, T un) {
    HCoef<T, VEC> c;
    auto ok = [&](int x) { return x >= 0 && x + 1 < W; };
#pragma unroll
    for (int k = 0; k < VEC; k += 2) {
        const bool v = ok(xl + k);
        c.ae[k / 2] = v ? a : un;
        c.be[k / 2] = v ? b : T(0);
    }
#pragma unroll
    for (int k = 1; k < VEC; k += 2) {
        const bool v = ok(xl + k) && (k + 1 < VEC || lane < kWave - 1);
        c.ao[k / 2] = v ? a : un;
        c.bo[k / 2] = v ? b : T(0);
    }
    const bool vl = ok(xl - 1) && lane > 0;
    c.al = vl ? a : un;
    c.bl = vl ? b : T(0);
    return c;
}

// H(s) in the fast form: every pair of the wave is inside the tile.
template <int PAR, bool IMAG, int MODE, typename T, int VEC>
__device__ __forceinline__ void hpass_fast(Row<T, VEC> &r, T k) {
    if constexpr (PAR == 0) {
#pragma unroll
        for (int j = 0; j < VEC; j += 2) fpair<IMAG, MODE>(k, r.re[j], r.im[j], r.re[j + 1], r.im[j + 1]);
    } else {
        const T nr = shfl_from_next(r.re[0]), ni = shfl_from_next(r.im[0]);
        const T pr = shfl_from_prev(r.re[VEC - 1]), pi = shfl_from_prev(r.im[VEC - 1]);
#pragma unroll
        for (int j = 1; j + 1 < VEC; j += 2) fpair<IMAG, MODE>(k, r.re[j], r.im[j], r.re[j + 1], r.im[j + 1]);
        fpair_p<IMAG, MODE>(k, r.re[VEC - 1], r.im[VEC - 1], nr, ni);
        fpair_q<IMAG, MODE>(k, pr, pi, r.re[0], r.im[0]);
    }
}

template <int PAR, bool IMAG, int MODE, typename T, int VEC>
__device__ __forceinline__ void vpair_fast(Row<T, VEC> &u, Row<T, VEC> &l, T k) {
    if constexpr (PAR == 0) {
#pragma unroll
        for (int j = 0; j < VEC; j += 2) fpair<IMAG, MODE>(k, u.re[j], u.im[j], l.re[j], l.im[j]);
    } else {
#pragma unroll
        for (int j = 1; j < VEC; j += 2) fpair<IMAG, MODE>(k, u.re[j], u.im[j], l.re[j], l.im[j]);
    }
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

// sin and cos of a float angle: three-part Cody-Waite reduction by pi/2 with
// fused multiply-adds (the reduced argument stays accurate to |x| ~ 2^22, far
// beyond any phase a Trotter step applies), then the Taylor polynomials of sin
// (to r^9) and cos (to r^10) on |r| <= pi/4, whose truncation errors (2e-9,
// 1e-10) are below half an ulp; the quadrant comes from the reduction integer.
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) {
    {
        const float k = rintf(x * 0.636619772367581343f);
        float r = fmaf(-k, 1.57079637050628662109375f, x);
        r = fmaf(-k, -4.37113900018624283e-8f, r);
        r = fmaf(-k, -1.71512451334011e-15f, r);
        const float r2 = r * r;
        float ps = fmaf(r2, 2.7557319e-6f, -1.98412698e-4f);
        ps = fmaf(r2, ps, 8.33333333e-3f);
        ps = fmaf(r2, ps, -1.66666667e-1f);
        const float sn = fmaf(r * r2, ps, r);
        float pc = fmaf(r2, -2.7557319e-7f, 2.48015873e-5f);
        pc = fmaf(r2, pc, -1.38888889e-3f);
        pc = fmaf(r2, pc, 4.16666667e-2f);
        pc = fmaf(r2, pc, -0.5f);
        const float cs = fmaf(r2, pc, 1.0f);
        const int q = (int)k;
        const float a = (q & 1) ? cs : sn;
        const float b = (q & 1) ? sn : cs;
        *s = (q & 2) ? -a : a;
        *c = ((q + 1) & 2) ? -b : b;
    }
}
__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ float exp_t(float x) { return expf(x); }
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }

// True when no lane of the wave holds a value above thr.
__device__ __forceinline__ bool wave_all_le(float v, float thr) {
    return __builtin_amdgcn_ballot_w64(!(v <= thr)) == 0;
}
__device__ __forceinline__ bool wave_all_le(double v, double thr) {
    return __builtin_amdgcn_ballot_w64(!(v <= thr)) == 0;
}

// sin/cos and exp(-x) of the nonlinear phase.  The phase g*dt*|psi|^2 of one
// step is small for any usable dt, so when every lane's |x| <= 1/64 (a
// wave-uniform test) short Taylor polynomials are already correctly rounded
// to within the format: truncation below 2e-10 relative in FP32 (sin to x^3,
// cos to x^4, exp to x^4) and below 1e-19 in FP64 (sin to x^7, cos to x^8,
// exp to x^8).  Otherwise the general routines run.
template <typename T, int VEC>
__device__ __forceinline__ void sincos_row(const T (&x)[VEC], T (&s)[VEC], T (&c)[VEC]) {
    T m = fabs(x[0]);
#pragma unroll
    for (int k = 1; k < VEC; ++k) m = fmax(m, fabs(x[k]));
    // the short polynomials always, the general routine over them in the rare
    // wave with a large angle (no copies on the common path)
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T r = x[k], r2 = r * r;
        if constexpr (sizeof(T) == 4) {
            s[k] = fmaf(r * r2, -0.166666667f, r);
            c[k] = fmaf(r2, fmaf(r2, 0.0416666667f, -0.5f), 1.0f);
        } else {
            T ps = fma(r2, -1.0 / 5040.0, 1.0 / 120.0);
            ps = fma(r2, ps, -1.0 / 6.0);
            s[k] = fma(r * r2, ps, r);
            T pc = fma(r2, 1.0 / 40320.0, -1.0 / 720.0);
            pc = fma(r2, pc, 1.0 / 24.0);
            pc = fma(r2, pc, -0.5);
            c[k] = fma(r2, pc, 1.0);
        }
    }
    if (__builtin_expect(!wave_all_le(m, T(0.015625)), 0)) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) sincos_t(x[k], &s[k], &c[k]);
    }
}

template <typename T, int VEC>
__device__ __forceinline__ void expneg_row(const T (&x)[VEC], T (&e)[VEC]) {
    T m = fabs(x[0]);
#pragma unroll
    for (int k = 1; k < VEC; ++k) m = fmax(m, fabs(x[k]));
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T r = -x[k];
        if constexpr (sizeof(T) == 4) {
            float q = fmaf(r, 0.0416666667f, 0.166666667f);
            q = fmaf(r, q, 0.5f);
            q = fmaf(r, q, 1.0f);
            e[k] = fmaf(r, q, 1.0f);
        } else {
            double q = fma(r, 1.0 / 40320.0, 1.0 / 5040.0);
            q = fma(r, q, 1.0 / 720.0);
            q = fma(r, q, 1.0 / 120.0);
            q = fma(r, q, 1.0 / 24.0);
            q = fma(r, q, 1.0 / 6.0);
            q = fma(r, q, 0.5);
            q = fma(r, q, 1.0);
            e[k] = fma(r, q, 1.0);
        }
    }
    if (__builtin_expect(!wave_all_le(m, T(0.015625)), 0)) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) e[k] = exp_t(-x[k]);
    }
}

// The phase-form rotation of a row, p -> mag e^(-i phi) p, from the rounded
// (cos, sin) of phi.  FP64: the factors cos*mag and sin*mag, then the rotation.
// FP32: k = mag (1 + e/2), e = 1 - cos^2 - sin^2, renormalises the rounded
// pair to first order (e from two fused multiply-adds, each square exact inside
// its fma), and mag = m1 + m2 is held in two floats so that no rounding of mag
// is applied uniformly.  k multiplies the state -- p k = fma(m1, p, u p),
// u = m2 + m1 e / 2, one rounding of a product that is inexact at almost every
// site -- and the raw (cos, sin) rotate it.  Forming cos*k and sin*k instead
// rounds them near the same values at every site of a slowly varying phase
// (where |phi| is small, cos rounds to 1 and cos*k to m1 everywhere): a bias of
// up to 1e-8 per step on the norm, against at most 1e-10 for p k (CPU
// emulation of both forms on random states and angles, |mag - 1| > 1e-6; within
// 1e-6 of 1, where k p rounds back to p, both leave up to 6e-9 at tiny angles).
template <typename T, int VEC>
__device__ __forceinline__ void rotate_row_ph(Row<T, VEC> &r, const T (&cs)[VEC], const T (&sn)[VEC],
                                              double mag) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        if constexpr (sizeof(T) == 4) {
            const float m1 = (float)mag;
            const float m2 = (float)(mag - (double)m1);
            float e = fmaf(-cs[k], cs[k], 1.0f);
            e = fmaf(-sn[k], sn[k], e);
            const float u = fmaf(0.5f * m1, e, m2);
            const float qr = fmaf(m1, r.re[k], u * r.re[k]);
            const float qi = fmaf(m1, r.im[k], u * r.im[k]);
            r.re[k] = fmaf(cs[k], qr, sn[k] * qi);
            r.im[k] = fmaf(cs[k], qi, -(sn[k] * qr));
        } else {
            const T c2 = cs[k] * mag, s2 = sn[k] * mag;
            const T pr = r.re[k], pi = r.im[k];
            r.re[k] = c2 * pr + s2 * pi;
            r.im[k] = c2 * pi - s2 * pr;
        }
    }
}

// P on one row (src/cpucartesian.cpp:93-109 / :127-137).  PH: e.re holds the
// phase phi (real time) or the exponent lambda (imaginary time) of E corr, which
// joins the nonlinear angle: one sincos (or exp) gives both factors, and |E|
// never differs from 1 by a rounding that repeats at the same site every step.
// The modulus |corr| is applied in FP64 before the factors are rounded.
template <bool IMAG, bool LHY, bool PH = false, typename T, int VEC>
__device__ __forceinline__ void ppass(Row<T, VEC> &r, const Row<T, VEC> &e, T c, T cl, double mag = 1.0) {
    T ang[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T n2 = r.re[k] * r.re[k] + r.im[k] * r.im[k];
        ang[k] = LHY ? c * n2 + cl * (n2 * sqrt_t(n2)) : c * n2;
        if (PH) ang[k] += e.re[k];
    }
    if (!IMAG) {
        T sn[VEC], cs[VEC];
        sincos_row<T, VEC>(ang, sn, cs);
        if (PH) {
            rotate_row_ph<T, VEC>(r, cs, sn, mag);
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                const T pr = r.re[k], pi = r.im[k];
                const T r1 = e.re[k] * pr - e.im[k] * pi;
                const T i1 = e.re[k] * pi + e.im[k] * pr;
                r.re[k] = cs[k] * r1 + sn[k] * i1;
                r.im[k] = cs[k] * i1 - sn[k] * r1;
            }
        }
    } else {
        T f[VEC];
        expneg_row<T, VEC>(ang, f);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const T g = PH ? f[k] : f[k] * e.re[k];
            r.re[k] = PH ? g * r.re[k] : f[k] * e.re[k] * r.re[k];
            r.im[k] = PH ? g * r.im[k] : f[k] * e.re[k] * r.im[k];
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

// Buffer-resource access: the base is the first row of the wave's segment, the
// lane offset is constant and the row offset is uniform, so addressing costs no
// vector instructions, and offsets outside the resource read as zero (stores are
// dropped) -- the lanes that must not touch memory get an offset past the end.
typedef unsigned __attribute__((ext_vector_type(4))) u32x4;
constexpr unsigned kOffNone = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

typedef unsigned __attribute__((ext_vector_type(2))) u32x2;

// A lane's VEC consecutive values: 16-byte pieces, or one 8-byte piece.
template <typename T, int VEC>
__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, T (&v)[VEC]) {
    static_assert(sizeof(T) * VEC % 16 == 0 || sizeof(T) * VEC == 8, "rows are loaded in 16- or 8-byte pieces");
    if constexpr (sizeof(T) * VEC == 8) {
        typedef T __attribute__((ext_vector_type(VEC))) TV;
        const TV x = __builtin_bit_cast(TV, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
#pragma unroll
        for (int k = 0; k < VEC; ++k) v[k] = x[k];
    } else {
        constexpr int P = sizeof(T) * VEC / 16;
        constexpr int E = 16 / sizeof(T);
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const u32x4 w = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * i, soff, 0));
            typedef T __attribute__((ext_vector_type(E))) TV;
            const TV x = __builtin_bit_cast(TV, w);
#pragma unroll
            for (int k = 0; k < E; ++k) v[i * E + k] = x[k];
        }
    }
}

// The row offset goes into the vector offset with a zero scalar offset: a
// 128-bit store then carries a one-wait-state hazard on its data registers
// that the compiler models (it does not for a store with a scalar-register
// offset).
// Store cache policy: kStoreNT marks the row non-temporal (streamed out of L2
// first), which leaves L2 to the input rows that neighbouring waves share.
constexpr int kStoreNT = 2;
// The host asks for it (template NT of step_kernel and step2_kernel) in FP64
// real time when a step's rows (state in and out, phase plane) exceed the
// 256 MB Infinity Cache: there the output is not read back from that cache, and
// the policy measured FP64 per step at 4096^2 87.7 -> 85.7 us, 8192^2 309.7 ->
// 302 us (two-step passes) and 595 -> 573 us (one-step); at 1024^2 and 2048^2,
// where the next step finds the rows in that cache, it measured slower (15.9
// -> 16.7 and 29.9 -> 31.0 us two-step, 18.2 -> 21.2 and 37.5 -> 44.5 us
// one-step).  FP32 keeps the default: its two-step pass measured 308.3 -> 313.0
// us with it at 8192^2 on two GPUs (305.4 on a third); imaginary time keeps it
// too (2048^2: 26.5 -> 29.1 us FP32, 43.8 -> 47.6 us FP64).
template <typename T, int VEC, int CPOL = 0>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const T (&v)[VEC]) {
    if constexpr (sizeof(T) * VEC == 8) {
        typedef T __attribute__((ext_vector_type(VEC))) TV;
        TV x;
#pragma unroll
        for (int k = 0; k < VEC; ++k) x[k] = v[k];
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, voff + soff, 0, CPOL);
    } else {
        constexpr int P = sizeof(T) * VEC / 16;
        constexpr int E = 16 / sizeof(T);
#pragma unroll
        for (int i = 0; i < P; ++i) {
            typedef T __attribute__((ext_vector_type(E))) TV;
            TV x;
#pragma unroll
            for (int k = 0; k < E; ++k) x[k] = v[i * E + k];
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, voff + soff + 16 * i, 0, CPOL);
        }
    }
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------
// One wavefront's share of a step: strip (wave mod n_strips), segment
// (wave / n_strips).  `wave` is wave-uniform.
// BLOCK_SUM (imaginary time): the wave returns its partial sum in *blk instead
// of storing it, and the caller combines the partials of its workgroup.
#ifndef TSMI_STEP_SYNC
#define TSMI_STEP_SYNC 4
#endif
// Beats between workgroup barriers in step_wave's all-fast loop (FP32; 0: none;
// a multiple of its 4 beats per iteration): the workgroup's four waves run
// neighbouring strips, which read the same apron columns, and in step they
// fetch those lines through L2 once.  8192^2 FP32 real time per launch: 308.7
// -> 299.8 us (every 12 and 24 beats: 301.0, 302.3 us); 2048^2 imaginary time
// 26.22 -> 25.88 us (the workgroup's closing barrier for the partial sums then
// follows the same number of barriers in each of its waves); FP64 real time
// 564.4 -> 566.9 us with it, so FP64 keeps none.
constexpr int kStepSync = TSMI_STEP_SYNC;
static_assert(kStepSync % 4 == 0, "the all-fast loop runs four beats per iteration");
constexpr int kStepBlock = 256;      // threads per step_kernel workgroup (its launch bound)

template <typename T, bool IMAG, int VEC, bool LHY, int MODE, bool PH = false, bool BLOCK_SUM = false, int SC = 0>
__device__ __forceinline__ void step_wave(const StepParams<T> &p, const int wave, const int lane,
                                          double *blk = nullptr, const int block_wave0 = -1) {
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

    const HCoef<T, VEC> hc = make_hcoef<T, VEC>(xl, lane, p.W, p.mH, p.tH, p.unH);
    const T mV = p.mV, tV = p.tV, unV = p.unV;
    const T kH = MODE == 0 ? p.tH : p.mH, kV = MODE == 0 ? p.tV : p.mV;
    // every horizontal pair of the wave lies inside the tile (MODE < 0: the two
    // axes use different scaled forms, so only the general form runs)
    const bool interior = MODE >= 0 && X0 >= 0 && X0 + kWave * VEC <= p.W;

    // Imaginary time: the pending renormalisation of the input, applied to each
    // row as it enters the pipeline (not as it is loaded, so that the loads of
    // the rows ahead stay in flight); computed after the first loads are issued.
    double divd = 1.0, dinv = 1.0;
    bool scale_in = false;

    double acc = 0.0, comp = 0.0;   // Neumaier sum of |psi|^2 over stored inner cells

    // segment-relative buffer resources
    const long seg0 = (long)y_lo * p.pitch;
    const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
    const __amdgpu_buffer_rsrc_t r_in_re = make_rsrc(p.in_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_in_im = make_rsrc(p.in_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_e_re = make_rsrc(PH ? p.ph + seg0 : p.e_re + seg0, seg_bytes);
    // the second E plane exists only for complex E (real time, E form)
    const __amdgpu_buffer_rsrc_t r_e_im = make_rsrc((IMAG || PH) ? (PH ? p.ph : p.e_re) + seg0 : p.e_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_out_re = make_rsrc(p.out_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_out_im = make_rsrc(p.out_im + seg0, seg_bytes);
    const unsigned row_bytes = (unsigned)p.pitch * (unsigned)sizeof(T);
    const unsigned v_ld = ld_ok ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    // own lanes store whole VEC-column pieces; columns in [W, pitch) are padding
    // and are written back as the zeros they hold
    const unsigned v_st = (own && ld_ok) ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    // inner columns of this lane for the imaginary-time sum
    unsigned inner_mask = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k)
        if (own && xl + k >= p.ix0 && xl + k < p.ix1) inner_mask |= 1u << k;

    auto load_psi = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_in_re, v_ld, so, r.re);
        buf_load<T, VEC>(r_in_im, v_ld, so, r.im);
    };
    auto load_e = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_e_re, v_ld, so, r.re);
        if (!IMAG && !PH) buf_load<T, VEC>(r_e_im, v_ld, so, r.im);
    };

    // Every row register starts at zero: rows outside the segment stay finite,
    // and a pair whose partner row does not exist gets the coefficients (1, 0).
    Row<T, VEC> pend1, pend2, pend3, pend4, inA, inB, eA, eB;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        pend1.re[k] = pend1.im[k] = pend2.re[k] = pend2.im[k] = T(0);
        pend3.re[k] = pend3.im[k] = pend4.re[k] = pend4.im[k] = T(0);
        inA.re[k] = inA.im[k] = inB.re[k] = inB.im[k] = T(0);
        eA.re[k] = eA.im[k] = eB.re[k] = eB.im[k] = T(0);
    }

    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };

    // Writes the row leaving the pipeline.
    // Imaginary time: adds the inner cells of an output row to the sum (the
    // lane's few cells of the row summed first, then one compensated step).
    // FP32: the lane's cells of the row are summed in FP32 (at most 2 VEC
    // products, a relative error below 1e-6 that the next renormalisation
    // removes), then join the FP64 sum.
    auto accumulate = [&](int y, const Row<T, VEC> &o4) {
        if (IMAG && y >= p.iy0 && y < p.iy1) {
            double v = 0.0;
            if constexpr (sizeof(T) == 4) {
                float v32 = 0.0f;
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const float w = (inner_mask & (1u << k)) ? o4.re[k] : 0.0f;
                    const float z = (inner_mask & (1u << k)) ? o4.im[k] : 0.0f;
                    v32 = fmaf(w, w, v32);
                    v32 = fmaf(z, z, v32);
                }
                v = double(v32);
            } else {
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const double vr = double(o4.re[k]), vi = double(o4.im[k]);
                    if (inner_mask & (1u << k)) v += vr * vr + vi * vi;
                }
            }
            const double t2 = acc + v;
            comp += (fabs(acc) >= fabs(v)) ? (acc - t2) + v : (v - t2) + acc;
            acc = t2;
        }
    };
    auto emit = [&](int y, const Row<T, VEC> &o4) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_store<T, VEC, SC>(r_out_re, v_st, so, o4.re);
        buf_store<T, VEC, SC>(r_out_im, v_st, so, o4.im);
        accumulate(y, o4);
    };
    // The emit of the fast form: a row outside [y_out_lo, y_out_hi) is dropped
    // by a store offset past the buffer, not by a branch.
    auto emit_any = [&](int y, const Row<T, VEC> &o4) {
        const bool out = y >= y_out_lo && y < y_out_hi;
        const unsigned off = out ? v_st + (unsigned)(y - y_lo) * row_bytes : kOffNone;
        buf_store<T, VEC, SC>(r_out_re, off, 0u, o4.re);
        buf_store<T, VEC, SC>(r_out_im, off, 0u, o4.im);
        if (out) accumulate(y, o4);
    };

    // One pipeline beat for input row t: stage k takes row t-k+1 and emits row
    // t-k.  TP = t mod 2 fixes every pair pattern (segments start on even rows,
    // see the host launcher).
    auto beat = [&](auto tp, int t, Row<T, VEC> &in, const Row<T, VEC> &e) {
        constexpr int TP = decltype(tp)::value;
        const bool v0 = valid(t), v1 = valid(t - 1), v2 = valid(t - 2), v3 = valid(t - 3), v4 = valid(t - 4);
        const T a1 = (v0 && v1) ? mV : unV, b1 = (v0 && v1) ? tV : T(0);
        const T a2 = (v1 && v2) ? mV : unV, b2 = (v1 && v2) ? tV : T(0);
        const T a3 = (v2 && v3) ? mV : unV, b3 = (v2 && v3) ? tV : T(0);
        const T a4 = (v3 && v4) ? mV : unV, b4 = (v3 && v4) ? tV : T(0);
        if (IMAG && scale_in && v0) scale_row<T, VEC>(in, dinv);
        // stage 1: V0, then H0 on row t-1
        vpair<(TP + 1) & 1, IMAG>(pend1, in, a1, b1);
        Row<T, VEC> o1 = pend1;
        pend1 = in;
        hpass<(TP + 1) & 1, IMAG>(o1, hc);
        // stage 2: V1, then H1 P H1 on row t-2
        vpair<(TP + 1) & 1, IMAG>(pend2, o1, a2, b2);
        Row<T, VEC> o2 = pend2;
        pend2 = o1;
        hpass<(TP + 1) & 1, IMAG>(o2, hc);
        ppass<IMAG, LHY, PH>(o2, e, p.c_nl, p.c_lhy, p.mag);
        hpass<(TP + 1) & 1, IMAG>(o2, hc);
        // stage 3: V1, then H0 on row t-3
        vpair<TP & 1, IMAG>(pend3, o2, a3, b3);
        Row<T, VEC> o3 = pend3;
        pend3 = o2;
        hpass<(TP + 1) & 1, IMAG>(o3, hc);
        // stage 4: V0; row t-4 leaves
        vpair<TP & 1, IMAG>(pend4, o3, a4, b4);
        const Row<T, VEC> o4 = pend4;
        pend4 = o3;
        const int y = t - 4;
        if (v4 && y >= y_out_lo && y < y_out_hi) emit(y, o4);
    };

    // The same beat in the fast form: interior strip, all five rows valid; the
    // row it emits is below the output range's end (steady state).
    auto beat_fast = [&](auto tp, int t, Row<T, VEC> &in, const Row<T, VEC> &e) {
        constexpr int TP = decltype(tp)::value;
        constexpr int M = MODE < 0 ? 0 : MODE;
        if (IMAG && scale_in) scale_row<T, VEC>(in, dinv);
        vpair_fast<(TP + 1) & 1, IMAG, M>(pend1, in, kV);
        Row<T, VEC> o1 = pend1;
        pend1 = in;
        hpass_fast<(TP + 1) & 1, IMAG, M>(o1, kH);
        vpair_fast<(TP + 1) & 1, IMAG, M>(pend2, o1, kV);
        Row<T, VEC> o2 = pend2;
        pend2 = o1;
        hpass_fast<(TP + 1) & 1, IMAG, M>(o2, kH);
        ppass<IMAG, LHY, PH>(o2, e, p.c_nl, p.c_lhy, p.mag);
        hpass_fast<(TP + 1) & 1, IMAG, M>(o2, kH);
        vpair_fast<TP & 1, IMAG, M>(pend3, o2, kV);
        Row<T, VEC> o3 = pend3;
        pend3 = o2;
        hpass_fast<(TP + 1) & 1, IMAG, M>(o3, kH);
        vpair_fast<TP & 1, IMAG, M>(pend4, o3, kV);
        const Row<T, VEC> o4 = pend4;
        pend4 = o3;
        emit_any(t - 4, o4);
    };

    // Software pipeline, four rows ahead: state rows t..t+3 and E rows t-2..t+1
    // are in flight while row t runs (a latency-bound wave -- small lattices,
    // short segments -- gains most).  A row that is not loaded (past the
    // segment) keeps the finite values of an older row; every pair touching it
    // has coefficients (1, 0).
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    int t = y_lo;
    const int t_end = y_hi + 4;
    Row<T, VEC> inC, inD, eC, eD;
#pragma unroll
    for (int k = 0; k < VEC; ++k) inC.re[k] = inC.im[k] = inD.re[k] = inD.im[k] = eC.re[k] = eC.im[k] = eD.re[k] = eD.im[k] = T(0);
    if (t < y_hi) load_psi(t, inA);
    if (t + 1 < y_hi) load_psi(t + 1, inB);
    if (t + 2 < y_hi) load_psi(t + 2, inC);
    if (t + 3 < y_hi) load_psi(t + 3, inD);
    if (t < y_hi) load_e(t, eA);
    if (t + 1 < y_hi) load_e(t + 1, eB);
    if (IMAG && p.part_in) {
        divd = renorm_divisor(p.part_in, __builtin_amdgcn_readfirstlane(*p.n_in), p.dxdy, p.norm_target, lane);
        dinv = 1.0 / divd;
        scale_in = divd != 1.0;
    }
    // slots: state row r in in[(r - y_lo) % 4]; E row r in e[(r - y_lo) % 4]
    auto group_general = [&](int t) {
        beat(I0(), t, inA, eC);
        if (t + 4 < y_hi) load_psi(t + 4, inA);
        if (t + 2 < y_hi) load_e(t + 2, eC);
        beat(I1(), t + 1, inB, eD);
        if (t + 5 < y_hi) load_psi(t + 5, inB);
        if (t + 3 < y_hi) load_e(t + 3, eD);
        beat(I0(), t + 2, inC, eA);
        if (t + 6 < y_hi) load_psi(t + 6, inC);
        if (t + 4 < y_hi) load_e(t + 4, eA);
        beat(I1(), t + 3, inD, eB);
        if (t + 7 < y_hi) load_psi(t + 7, inD);
        if (t + 5 < y_hi) load_e(t + 5, eB);
    };
    // The steady state of an interior strip (t >= y_lo + 4, t + 3 < y_hi), in a
    // loop of its own with no conditional memory operation: the single loop
    // merged fast and general paths with different counts of outstanding
    // operations, where the compiler drained them, and needed more registers
    // (FP32: 204 VGPRs against 161).  A row past the segment is addressed
    // through the vector offset, which the buffer's range check turns into a
    // load of zeros (such a row only meets pairs with coefficients (1, 0)), and
    // a row before the output range (the first four beats) gets a store offset
    // past the buffer, which drops it.
    auto load_psi_any = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_in_re, v_ld + so, 0u, r.re);
        buf_load<T, VEC>(r_in_im, v_ld + so, 0u, r.im);
    };
    auto load_e_any = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_e_re, v_ld + so, 0u, r.re);
        if (!IMAG && !PH) buf_load<T, VEC>(r_e_im, v_ld + so, 0u, r.im);
    };
    auto group_fast = [&](int t) {
        beat_fast(I0(), t, inA, eC);
        load_psi_any(t + 4, inA);
        load_e_any(t + 2, eC);
        beat_fast(I1(), t + 1, inB, eD);
        load_psi_any(t + 5, inB);
        load_e_any(t + 3, eD);
        beat_fast(I0(), t + 2, inC, eA);
        load_psi_any(t + 6, inC);
        load_e_any(t + 4, eA);
        beat_fast(I1(), t + 3, inD, eB);
        load_psi_any(t + 7, inD);
        load_e_any(t + 5, eB);
    };
    // Every beat in the fast form when the segment's four apron rows above and
    // below lie inside the lattice: the zero rows the pipeline starts with and
    // the zero rows past the segment then meet apron rows only.  A pair with a
    // wrong partner leaves a wrong value in an apron row; one vertical pass
    // moves it at most one row, four a step, so it never reaches an output row
    // (the 4-column strip aprons rest on the same bound).  The row offsets of
    // the last beats stay below 2^31.
    const bool rows_ok = (unsigned long)(p.seg_rows + 2 * kApron + 12) * row_bytes < 0x80000000ul;
    auto all_fast_of = [&](int w) {
        const int st = w % p.n_strips, sg = w / p.n_strips;
        const int x0 = st * SW - kApron, lo = sg * p.seg_rows;
        return w < p.n_waves && MODE >= 0 && x0 >= 0 && x0 + kWave * VEC <= p.W && lo >= kApron &&
               min(p.H, lo + p.seg_rows) + kApron <= p.H && rows_ok;
    };
    const bool all_fast = all_fast_of(wave);
    if (all_fast) {
        // real time, all four waves of the workgroup here (the same segment
        // height, so the same trip count): a barrier every kStepSync beats
        bool sync = false;
        if constexpr (sizeof(T) == 4 && kStepSync > 0) {
            sync = block_wave0 >= 0;
#pragma unroll
            for (int i = 0; i < kStepBlock / kWave; ++i) sync = sync && all_fast_of(block_wave0 + i);
        }
        for (int it = 0; t < t_end; t += 4, ++it) {
            group_fast(t);
            if constexpr (sizeof(T) == 4 && kStepSync > 0) {
                if (sync && it % (kStepSync / 4) == kStepSync / 4 - 1) __builtin_amdgcn_s_barrier();
            }
        }
    } else if (interior) {
        // rows y_lo .. y_lo+3: the pipeline fills
        for (; t < t_end && t < y_lo + 4; t += 4) group_general(t);
        for (; t + 3 < y_hi; t += 4) group_fast(t);
    }
    for (; t < t_end; t += 4) group_general(t);

    if (IMAG) {
        // this wave's partial sum for the next step's divisor
        double s = acc + comp;
        for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
        if constexpr (BLOCK_SUM) {
            *blk = s;
        } else if (lane == 0) {
            p.part_out[wave] = s;
            if (wave == 0) *p.n_out = p.n_waves;
        }
    }
}

// One launch of one step.  Imaginary time: the partial sums of a workgroup's
// waves are combined in wave order into one partial per workgroup (LDS and one
// barrier at the end), so the next launch's waves each read a quarter as many
// partials (the read is O(waves x partials) per step).  Every wave of the grid
// reaches that barrier, those past the last segment included.
template <typename T, bool IMAG, int VEC, bool LHY, int MODE, bool PH = false, bool NT = false>
__global__ void __launch_bounds__(kStepBlock) step_kernel(StepParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform by construction; readfirstlane lets the compiler keep every
    // quantity derived from it (segment, strip, row bounds) in scalar registers
    const int wave0 = xcd_block<IMAG || sizeof(T) == 8>(blockIdx.x, gridDim.x) * (blockDim.x / kWave);
    const int wave = __builtin_amdgcn_readfirstlane(wave0 + (threadIdx.x / kWave));
    if constexpr (!IMAG) {
        if (wave >= p.n_waves) return;
        step_wave<T, IMAG, VEC, LHY, MODE, PH, false, NT ? kStoreNT : 0>(p, wave, lane, nullptr, wave0);
    } else {
        constexpr int kWavesPerBlock = 256 / kWave;
        __shared__ double part[kWavesPerBlock];
        double s = 0.0;
        if (wave < p.n_waves) step_wave<T, IMAG, VEC, LHY, MODE, PH, true>(p, wave, lane, &s, wave0);
        if (lane == 0) part[threadIdx.x / kWave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, c = 0.0;
#pragma unroll
            for (int i = 0; i < kWavesPerBlock; ++i) {
                const double v = part[i];
                const double t = a + v;
                c += (fabs(a) >= fabs(v)) ? (a - t) + v : (v - t) + a;
                a = t;
            }
            p.part_out[blockIdx.x] = a + c;
            if (blockIdx.x == 0) *p.n_out = (int)gridDim.x;
        }
    }
}

// ---------------------------------------------------------------------------
// periodic halo exchange
// ---------------------------------------------------------------------------
// The single-process exchange of the reference (src/cpukernel.cpp:707-715,
// then :738-746): first the x halo of rows [rx0, rx0 + H - 2*halo_y) (that row
// count holds even on a closed y axis) is copied from the opposite inner edge,
// then the full-width y halo rows.  Every destination is computed from its final
// source directly (a y-halo cell whose source row had its x halo refreshed
// takes the inner cell both copies lead to), so sources are never destinations
// and one launch suffices.
// Destinations: [0, hy) and [H-hy, H) full rows (periodic y), then the x bands
// of the rows in between (periodic x).
__host__ __device__ __forceinline__ int exchange_cells(int W, int H, int px, int py, int hx, int hy) {
    return (py ? 2 * hy * W : 0) + (px ? 2 * hx * (H - (py ? 2 * hy : 0)) : 0);
}

// Refreshes destination cell i (0 <= i < exchange_cells(...)).
template <typename T>
__device__ __forceinline__ void exchange_cell(int i, T *re, T *im, int W, int H, int pitch, int px, int py,
                                              int hx, int hy) {
    const int nx = W - 2 * hx, ny = H - 2 * hy;
    const int rx0 = py ? hy : 0, rx1 = rx0 + max(0, H - 2 * hy);
    const int n_yband = py ? 2 * hy * W : 0;
    int x, y;
    if (i < n_yband) {
        const int r = i / W;
        x = i - r * W;
        y = r < hy ? r : H - 2 * hy + r;
    } else {
        const int j = i - n_yband;
        const int r = j / (2 * hx);
        const int c = j - r * 2 * hx;
        y = (py ? hy : 0) + r;
        x = c < hx ? c : W - 2 * hx + c;
    }
    int sy = y, sx = x;
    if (py && (y < hy || y >= H - hy)) sy = y < hy ? y + ny : y - ny;
    const bool xd = px && (x < hx || x >= W - hx) && sy >= rx0 && sy < rx1;
    if (xd) sx = x < hx ? x + nx : x - nx;
    if (sx == x && sy == y) return;
    const long d = (long)y * pitch + x, s = (long)sy * pitch + sx;
    re[d] = re[s];
    im[d] = im[s];
}

template <typename T>
__global__ void __launch_bounds__(256) halo_exchange_kernel(T *re, T *im, int W, int H, int pitch, int px, int py,
                                                            int hx, int hy) {
    const int n = exchange_cells(W, H, px, py, hx, hy);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        exchange_cell<T>(i, re, im, W, H, pitch, px, py, hx, hy);
}

// The same exchange on a doubly periodic tile in 16-byte pieces (4 FP32 or 2
// FP64 cells), when hx, W and the pitch are multiples of a piece, so that no
// piece straddles a halo edge: items [0, n_x) copy the x halo of rows
// [hy, H - hy) (row r, side s, piece k), items [n_x, n_x + n_y) the full-width
// y halo rows, each piece from its wrapped row and, in the x halo, its wrapped
// columns -- the final source of every cell, as in exchange_cell.
__host__ __device__ __forceinline__ int exchange_pieces(int W, int H, int hx, int hy, int v) {
    return (H - 2 * hy) * 2 * (hx / v) + 2 * hy * (W / v);
}

template <typename T>
__global__ void __launch_bounds__(256) halo_exchange_vec_kernel(T *re, T *im, int W, int H, int pitch, int hx,
                                                                int hy) {
    typedef T __attribute__((ext_vector_type(16 / sizeof(T)))) TV;
    constexpr int V = 16 / sizeof(T);
    const int nx = W - 2 * hx, ny = H - 2 * hy;
    const int kx = hx / V, n_x = (H - 2 * hy) * 2 * kx, wv = W / V;
    const int n = n_x + 2 * hy * wv;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int x, y, sx, sy;
        if (i < n_x) {
            const int r = i / (2 * kx), j = i - r * 2 * kx;
            y = hy + r;
            sy = y;
            x = j < kx ? j * V : W - hx + (j - kx) * V;
            sx = j < kx ? x + nx : x - nx;
        } else {
            const int j = i - n_x, r = j / wv;
            x = (j - r * wv) * V;
            y = r < hy ? r : H - 2 * hy + r;
            sy = r < hy ? y + ny : y - ny;
            sx = x < hx ? x + nx : (x >= W - hx ? x - nx : x);
        }
        const long d = (long)y * pitch + x, s = (long)sy * pitch + sx;
        *(TV *)(re + d) = *(const TV *)(re + s);
        *(TV *)(im + d) = *(const TV *)(im + s);
    }
}

// The transpose of the exchange on a field of derivatives (G_src += G_dst,
// G_dst = 0 for every copy, the exchange's two stages in reverse order): stage
// 0 folds the y-halo rows (full width), stage 1 the x-halo columns of the rows
// the exchange refreshes.  Each source cell has one destination per stage when
// the inner extent is at least twice the halo on a periodic axis.
__host__ __device__ __forceinline__ int fold_cells(int W, int H, int px, int py, int hx, int hy, int stage) {
    if (stage == 0) return py ? 2 * hy * W : 0;
    return px ? 2 * hx * max(0, H - 2 * hy) : 0;
}

template <typename T>
__global__ void __launch_bounds__(256) halo_fold_kernel(T *re, T *im, int W, int H, int pitch, int px, int py,
                                                        int hx, int hy, int stage) {
    const int n = fold_cells(W, H, px, py, hx, hy, stage);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        long d, s;
        if (stage == 0) {
            const int r = i / W, x = i - r * W;
            const int dy = r < hy ? r : H - 2 * hy + r;
            const int sy = r < hy ? dy + (H - 2 * hy) : dy - (H - 2 * hy);
            d = (long)dy * pitch + x;
            s = (long)sy * pitch + x;
        } else {
            const int r = i / (2 * hx), c = i - r * 2 * hx;
            const int y = (py ? hy : 0) + r;
            const int dx = c < hx ? c : W - 2 * hx + c;
            const int sx = c < hx ? dx + (W - 2 * hx) : dx - (W - 2 * hx);
            d = (long)y * pitch + dx;
            s = (long)y * pitch + sx;
        }
        re[s] += re[d];
        im[s] += im[d];
        re[d] = T(0);
        im[d] = T(0);
    }
}

// Zeros the cells of a W x H field outside the inner region [x0, x1) x [y0, y1).
template <typename T>
__global__ void __launch_bounds__(256) zero_outside_kernel(T *p, int W, int H, int pitch, int x0, int x1, int y0,
                                                           int y1) {
    const long n = (long)W * H;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long)y * W);
        if (x < x0 || x >= x1 || y < y0 || y >= y1) p[(long)y * pitch + x] = T(0);
    }
}

}  // namespace tsmi
