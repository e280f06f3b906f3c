// pair_kernel.hpp -- both components of a two-component lattice, one step per
// pass (pair_kernel) or, in FP32 real time, two (pair2_kernel).
//
// The reference advances a two-component system (Hamiltonian2Component with the
// two-state CPUBlock, src/cpukernel.cpp:274-392) by running the single-component
// step once per component (src/solver.cpp:201-221): component a's potential pass
// sees the other component's density |psi_b|^2, component b's sees |psi_a|^2
// (src/cpucartesian.cpp:75-92 real time, :115-126 imaginary time), and both
// densities are the start-of-step ones -- b's step reads a's old buffer
// (src/cpukernel.cpp:421-424 after wait_for_completion flips `sense` back,
// :565-569).  The step ends with the Rabi rotation that mixes the two
// (src/cpukernel.cpp:581-608, src/cpucartesian.cpp:264-290).
//
// Since neither component's step reads the other's new values, one wavefront
// runs both row pipelines of its strip segment side by side (pair_wave): each
// input row's start-of-step density stays in registers for the two beats until
// the potential pass of that row, so the other component costs no extra HBM
// traffic.  In real time the Rabi rotation is applied to each pair of rows as it
// leaves the pipelines; per site a step reads psi_a, psi_b, E_a, E_b once and
// writes psi_a, psi_b once.  In imaginary time each component's renormalisation
// (a global sum) comes before the rotation, so the rotation and the joint
// renormalisation run as their own pass (rabi_kernel) when the coupling is on.
#pragma once
#include "step_kernel.hpp"

namespace tsmi {

template <typename T>
struct PairParams {
    const T *in_re[2], *in_im[2];
    T *out_re[2], *out_im[2];
    const T *e_re[2], *e_im[2];      // E of each component (e_im unused in imaginary time)
    const T *ph[2];                  // one-plane form of each E (see StepParams)
    double mag[2];                   // |corr| of each component (real time, phase form)
    int W, H, pitch;
    int n_strips, n_waves, seg_rows;
    T mH[2], tH[2], mV[2], tV[2], unH[2], unV[2];   // scaled-frame pair factors per component
    T c_self[2];                     // dt g of each component, in its mid-step frame
    T c_ab;                          // dt g_ab (the other density is a stored, true-frame value)
    // Real-time Rabi rotation of the rows leaving the pipelines (constants in
    // FP64; R = [[c, i s], [i conj(s), c]] has R^H R = (c^2 + |s|^2) I, so an FP32
    // c rounded near 1 would scale the norm by the same factor every step -- the
    // FP32 kernels use c = 1 - d, see rabi_rows).
    double cc, cs_r, cs_i;
    double cc2, cs_r2, cs_i2;        // the second step's rotation in a two-step pass (pair2_kernel)
    int rabi;                        // 1: apply it
    // imaginary time
    const double *div_in[2];         // pending divisors of the inputs (nullptr: 1)
    double *div_out;                 // divisors of this step's outputs, one per component
    double *partials;                // [2][n_waves]
    unsigned *counter;               // zero between launches
    const double *norm;              // per-component squared-norm targets (0: no renormalisation)
    double dxdy;
    int ix0, ix1, iy0, iy1;
};

// P of one component (src/cpucartesian.cpp:75-92 / :115-126): the phase or decay
// sees its own density (scaled frame, c already rescaled) and the other
// component's start-of-step density nb.
template <bool IMAG, bool PH, typename T, int VEC>
__device__ __forceinline__ void ppass_pair(Row<T, VEC> &r, const Row<T, VEC> &e, const T (&nb)[VEC], T c, T cab,
                                           double mag) {
    T ang[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T n2 = r.re[k] * r.re[k] + r.im[k] * r.im[k];
        ang[k] = c * n2 + cab * nb[k];
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
            r.re[k] = g * r.re[k];
            r.im[k] = g * r.im[k];
        }
    }
}

template <typename T, int VEC>
struct Pipe4 {
    Row<T, VEC> p1, p2, p3, p4;
};

// One beat of one component's pipeline (the general form of step_wave's beat):
// input row t enters, row t-4 leaves in o4.
template <int TP, bool IMAG, bool PH, typename T, int VEC>
__device__ __forceinline__ void pair_pipe(Pipe4<T, VEC> &q, Row<T, VEC> &in, const Row<T, VEC> &e,
                                          const T (&nb)[VEC], Row<T, VEC> &o4, bool v0, bool v1, bool v2, bool v3,
                                          bool v4, const HCoef<T, VEC> &hc, T mV, T tV, T unV, T c, T cab,
                                          double mag) {
    const T a1 = (v0 && v1) ? mV : unV, b1 = (v0 && v1) ? tV : T(0);
    const T a2 = (v1 && v2) ? mV : unV, b2 = (v1 && v2) ? tV : T(0);
    const T a3 = (v2 && v3) ? mV : unV, b3 = (v2 && v3) ? tV : T(0);
    const T a4 = (v3 && v4) ? mV : unV, b4 = (v3 && v4) ? tV : T(0);
    // stage 1: V0, then H0 on row t-1
    vpair<(TP + 1) & 1, IMAG>(q.p1, in, a1, b1);
    Row<T, VEC> o1 = q.p1;
    q.p1 = in;
    hpass<(TP + 1) & 1, IMAG>(o1, hc);
    // stage 2: V1, then H1 P H1 on row t-2
    vpair<(TP + 1) & 1, IMAG>(q.p2, o1, a2, b2);
    Row<T, VEC> o2 = q.p2;
    q.p2 = o1;
    hpass<(TP + 1) & 1, IMAG>(o2, hc);
    ppass_pair<IMAG, PH>(o2, e, nb, c, cab, mag);
    hpass<(TP + 1) & 1, IMAG>(o2, hc);
    // stage 3: V1, then H0 on row t-3
    vpair<TP & 1, IMAG>(q.p3, o2, a3, b3);
    Row<T, VEC> o3 = q.p3;
    q.p3 = o2;
    hpass<(TP + 1) & 1, IMAG>(o3, hc);
    // stage 4: V0; row t-4 leaves
    vpair<TP & 1, IMAG>(q.p4, o3, a4, b4);
    o4 = q.p4;
    q.p4 = o3;
}

// The same beat in the fast scaled form (one fused multiply-add per output, as
// step_wave's fast beat): every pair of the beat inside the tile, both
// components in form MODE on both axes.
template <int TP, int MODE, bool PH, typename T, int VEC>
__device__ __forceinline__ void pair_pipe_fast(Pipe4<T, VEC> &q, Row<T, VEC> &in, const Row<T, VEC> &e,
                                               const T (&nb)[VEC], Row<T, VEC> &o4, T kH, T kV, T c, T cab,
                                               double mag) {
    vpair_fast<(TP + 1) & 1, false, MODE>(q.p1, in, kV);
    Row<T, VEC> o1 = q.p1;
    q.p1 = in;
    hpass_fast<(TP + 1) & 1, false, MODE>(o1, kH);
    vpair_fast<(TP + 1) & 1, false, MODE>(q.p2, o1, kV);
    Row<T, VEC> o2 = q.p2;
    q.p2 = o1;
    hpass_fast<(TP + 1) & 1, false, MODE>(o2, kH);
    ppass_pair<false, PH>(o2, e, nb, c, cab, mag);
    hpass_fast<(TP + 1) & 1, false, MODE>(o2, kH);
    vpair_fast<TP & 1, false, MODE>(q.p3, o2, kV);
    Row<T, VEC> o3 = q.p3;
    q.p3 = o2;
    hpass_fast<(TP + 1) & 1, false, MODE>(o3, kH);
    vpair_fast<TP & 1, false, MODE>(q.p4, o3, kV);
    o4 = q.p4;
    q.p4 = o3;
}

// The real-time Rabi rotation of a pair of rows (src/cpucartesian.cpp:264-276,
// every right-hand side the old value).  FP64 in the FP64 kernels.  FP32 with
// c = 1 - d: each output is x plus the increment -d x + (the two s terms),
// formed in FP32 first and added to x in one rounding.  R^H R = ((1 - d)^2 +
// |s|^2) I then differs from I only by the roundings of d and s relative to
// themselves (about 1e-7 of |s|^2 ~ 2d).  An FP32 c rounded near 1 would put a
// factor off by up to 6e-8 on the norm every step, and so would -d x applied in
// a rounding of its own: d (3e-8 at Omega dt = 5e-4) is below half an ulp of x,
// so x - d x rounds back to x at most sites while the s terms still add |s|^2.
template <typename T, int VEC>
__device__ __forceinline__ void rabi_rows(Row<T, VEC> &oa, Row<T, VEC> &ob, double cc, double sr, double si) {
    if constexpr (sizeof(T) == 4) {
        const float d = float(1.0 - cc), fr = float(sr), fi = float(si);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const float ar = oa.re[k], ai = oa.im[k], br = ob.re[k], bi = ob.im[k];
            const float ua = fmaf(-d, ar, fmaf(-fi, br, -(fr * bi)));
            const float va = fmaf(-d, ai, fmaf(fr, br, -(fi * bi)));
            const float ub = fmaf(-d, br, fmaf(fi, ar, -(fr * ai)));
            const float vb = fmaf(-d, bi, fmaf(fr, ar, fi * ai));
            oa.re[k] = ar + ua;
            oa.im[k] = ai + va;
            ob.re[k] = br + ub;
            ob.im[k] = bi + vb;
        }
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const double ar = oa.re[k], ai = oa.im[k], br = ob.re[k], bi = ob.im[k];
            oa.re[k] = T(cc * ar - si * br - sr * bi);
            oa.im[k] = T(cc * ai + sr * br - si * bi);
            ob.re[k] = T(cc * br + si * ar - sr * ai);
            ob.im[k] = T(cc * bi + sr * ar + si * ai);
        }
    }
}

template <typename T, int VEC>
__device__ __forceinline__ void density(const Row<T, VEC> &r, T (&n)[VEC]) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) n[k] = r.re[k] * r.re[k] + r.im[k] * r.im[k];
}

__device__ __forceinline__ void neumaier(double &s, double &c, double v) {
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
}

template <typename T, bool IMAG, int VEC, bool PH, int MODE>
__device__ __forceinline__ void pair_wave(const PairParams<T> &p, const int wave, const int lane) {
    constexpr int SW = kWave * VEC - 2 * kApron;
    const int strip = wave % p.n_strips;
    const int seg = wave / p.n_strips;
    const int X0 = strip * SW - kApron;
    const int xl = X0 + lane * VEC;
    const int y_out_lo = seg * p.seg_rows;
    const int y_out_hi = min(p.H, y_out_lo + p.seg_rows);
    const int y_lo = max(0, y_out_lo - kApron);
    const int y_hi = min(p.H, y_out_hi + kApron);
    const bool ld_ok = xl >= 0 && xl + VEC <= p.pitch;
    const bool own = lane >= kApron / VEC && lane < kWave - kApron / VEC;

    const HCoef<T, VEC> hc0 = make_hcoef<T, VEC>(xl, lane, p.W, p.mH[0], p.tH[0], p.unH[0]);
    const HCoef<T, VEC> hc1 = make_hcoef<T, VEC>(xl, lane, p.W, p.mH[1], p.tH[1], p.unH[1]);
    const T mV0 = p.mV[0], tV0 = p.tV[0], unV0 = p.unV[0], c0 = p.c_self[0];
    const T mV1 = p.mV[1], tV1 = p.tV[1], unV1 = p.unV[1], c1 = p.c_self[1];
    const T cab = p.c_ab;

    const bool div0_on = IMAG && p.div_in[0] != nullptr, div1_on = IMAG && p.div_in[1] != nullptr;
    // pending divisors, applied to each row as it enters the pipelines (not as
    // it is loaded, so that the loads of the rows ahead stay in flight)
    const double div0 = div0_on ? *p.div_in[0] : 1.0, div1 = div1_on ? *p.div_in[1] : 1.0;
    const double inv0 = 1.0 / div0, inv1 = 1.0 / div1;
    double acc0 = 0.0, cmp0 = 0.0, acc1 = 0.0, cmp1 = 0.0;

    const long seg0 = (long)y_lo * p.pitch;
    const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
    const __amdgpu_buffer_rsrc_t ri0r = make_rsrc(p.in_re[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri0i = make_rsrc(p.in_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri1r = make_rsrc(p.in_re[1] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri1i = make_rsrc(p.in_im[1] + seg0, seg_bytes);
    // E of each component, or its one-plane form (the second plane exists only
    // for complex E in the E form)
    const T *e0 = PH ? p.ph[0] : p.e_re[0], *e1 = PH ? p.ph[1] : p.e_re[1];
    const __amdgpu_buffer_rsrc_t re0r = make_rsrc(e0 + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re0i = make_rsrc((IMAG || PH) ? e0 + seg0 : p.e_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re1r = make_rsrc(e1 + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re1i = make_rsrc((IMAG || PH) ? e1 + seg0 : p.e_im[1] + seg0, seg_bytes);
    const double mag0 = p.mag[0], mag1 = p.mag[1];
    const __amdgpu_buffer_rsrc_t ro0r = make_rsrc(p.out_re[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro0i = make_rsrc(p.out_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro1r = make_rsrc(p.out_re[1] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro1i = make_rsrc(p.out_im[1] + seg0, seg_bytes);
    const unsigned row_bytes = (unsigned)p.pitch * (unsigned)sizeof(T);
    const unsigned v_ld = ld_ok ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    const unsigned v_st = (own && ld_ok) ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    unsigned inner_mask = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k)
        if (own && xl + k >= p.ix0 && xl + k < p.ix1) inner_mask |= 1u << k;

    auto load_psi = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(ri0r, v_ld, so, a.re);
        buf_load<T, VEC>(ri0i, v_ld, so, a.im);
        buf_load<T, VEC>(ri1r, v_ld, so, b.re);
        buf_load<T, VEC>(ri1i, v_ld, so, b.im);
    };
    auto load_e = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(re0r, v_ld, so, a.re);
        buf_load<T, VEC>(re1r, v_ld, so, b.re);
        if (!IMAG && !PH) {
            buf_load<T, VEC>(re0i, v_ld, so, a.im);
            buf_load<T, VEC>(re1i, v_ld, so, b.im);
        }
    };

    Pipe4<T, VEC> qa, qb;
    Row<T, VEC> ia0, ib0, ia1, ib1, ea0, eb0, ea1, eb1;
    T na0[VEC], nb0[VEC], na1[VEC], nb1[VEC];   // start-of-step densities, slot by row parity
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        qa.p1.re[k] = qa.p1.im[k] = qa.p2.re[k] = qa.p2.im[k] = T(0);
        qa.p3.re[k] = qa.p3.im[k] = qa.p4.re[k] = qa.p4.im[k] = T(0);
        qb.p1.re[k] = qb.p1.im[k] = qb.p2.re[k] = qb.p2.im[k] = T(0);
        qb.p3.re[k] = qb.p3.im[k] = qb.p4.re[k] = qb.p4.im[k] = T(0);
        ia0.re[k] = ia0.im[k] = ib0.re[k] = ib0.im[k] = ia1.re[k] = ia1.im[k] = ib1.re[k] = ib1.im[k] = T(0);
        ea0.re[k] = ea0.im[k] = eb0.re[k] = eb0.im[k] = ea1.re[k] = ea1.im[k] = eb1.re[k] = eb1.im[k] = T(0);
        na0[k] = nb0[k] = na1[k] = nb1[k] = T(0);
    }
    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };

    auto emit = [&](int y, Row<T, VEC> &oa, Row<T, VEC> &ob) {
        if (!IMAG && p.rabi) rabi_rows<T, VEC>(oa, ob, p.cc, p.cs_r, p.cs_i);
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_store<T, VEC>(ro0r, v_st, so, oa.re);
        buf_store<T, VEC>(ro0i, v_st, so, oa.im);
        buf_store<T, VEC>(ro1r, v_st, so, ob.re);
        buf_store<T, VEC>(ro1i, v_st, so, ob.im);
        if (IMAG && y >= p.iy0 && y < p.iy1) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                if (inner_mask & (1u << k)) {
                    const double ar = double(oa.re[k]), ai = double(oa.im[k]);
                    const double br = double(ob.re[k]), bi = double(ob.im[k]);
                    neumaier(acc0, cmp0, ar * ar + ai * ai);
                    neumaier(acc1, cmp1, br * br + bi * bi);
                }
            }
        }
    };

    // beat t: both components take input row t; their potential passes (row
    // t-2) see the other's density of row t-2, which slot n[a|b] holds until it
    // is replaced by row t's.
    auto beat = [&](auto tp, int t, Row<T, VEC> &ia, Row<T, VEC> &ib, const Row<T, VEC> &ea,
                    const Row<T, VEC> &eb, T (&na)[VEC], T (&nb)[VEC]) {
        constexpr int TP = decltype(tp)::value;
        const bool v0 = valid(t), v1 = valid(t - 1), v2 = valid(t - 2), v3 = valid(t - 3), v4 = valid(t - 4);
        if (div0_on && v0) scale_row<T, VEC>(ia, inv0);
        if (div1_on && v0) scale_row<T, VEC>(ib, inv1);
        T ta[VEC], tb[VEC];
        density(ia, ta);
        density(ib, tb);
        Row<T, VEC> oa, ob;
        pair_pipe<TP, IMAG, PH>(qa, ia, ea, nb, oa, v0, v1, v2, v3, v4, hc0, mV0, tV0, unV0, c0, cab, mag0);
        pair_pipe<TP, IMAG, PH>(qb, ib, eb, na, ob, v0, v1, v2, v3, v4, hc1, mV1, tV1, unV1, c1, cab, mag1);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            na[k] = ta[k];
            nb[k] = tb[k];
        }
        const int y = t - 4;
        if (v4 && y >= y_out_lo && y < y_out_hi) emit(y, oa, ob);
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    int t = y_lo;
    if (t < y_hi) load_psi(t, ia0, ib0);
    if (t + 1 < y_hi) load_psi(t + 1, ia1, ib1);
    const int t_end = y_hi + 4;
    // FP64 real time: every beat in the fast form when the strip lies inside the
    // tile and the segment's 4 apron rows above and below inside the lattice (the
    // zero rows the pipelines start with and the zero rows loaded past the
    // segment meet apron rows only; a wrong value moves at most one row per
    // vertical pass, four per step).  4096^2: 0.321 -> 0.304 ms per step; in FP32
    // (3 waves per SIMD) the same path measured slower, 0.170 -> 0.197 ms.
    if constexpr (!IMAG && MODE >= 0 && sizeof(T) == 8) {
        if (X0 >= 0 && X0 + kWave * VEC <= p.W && y_out_lo >= kApron && y_out_hi + kApron <= p.H) {
            const T kH0 = MODE == 0 ? p.tH[0] : p.mH[0], kV0 = MODE == 0 ? p.tV[0] : p.mV[0];
            const T kH1 = MODE == 0 ? p.tH[1] : p.mH[1], kV1 = MODE == 0 ? p.tV[1] : p.mV[1];
            auto beat_fast = [&](auto tp, int t, Row<T, VEC> &ia, Row<T, VEC> &ib, const Row<T, VEC> &ea,
                                 const Row<T, VEC> &eb, T (&na)[VEC], T (&nb)[VEC]) {
                constexpr int TP = decltype(tp)::value;
                T ta[VEC], tb[VEC];
                density(ia, ta);
                density(ib, tb);
                Row<T, VEC> oa, ob;
                pair_pipe_fast<TP, MODE, PH>(qa, ia, ea, nb, oa, kH0, kV0, c0, cab, mag0);
                pair_pipe_fast<TP, MODE, PH>(qb, ib, eb, na, ob, kH1, kV1, c1, cab, mag1);
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    na[k] = ta[k];
                    nb[k] = tb[k];
                }
                const int y = t - 4;
                if (y >= y_out_lo && y < y_out_hi) emit(y, oa, ob);
            };
            // rows past the segment load as zeros (the buffer range check)
            for (; t < t_end; t += 2) {
                beat_fast(I0(), t, ia0, ib0, ea0, eb0, na0, nb0);
                load_psi(t + 2, ia0, ib0);
                load_e(t, ea0, eb0);
                beat_fast(I1(), t + 1, ia1, ib1, ea1, eb1, na1, nb1);
                load_psi(t + 3, ia1, ib1);
                load_e(t + 1, ea1, eb1);
            }
            return;
        }
    }
    auto pair_general = [&](int t) {
        beat(I0(), t, ia0, ib0, ea0, eb0, na0, nb0);
        if (t + 2 < y_hi) load_psi(t + 2, ia0, ib0);
        if (t < y_hi) load_e(t, ea0, eb0);
        beat(I1(), t + 1, ia1, ib1, ea1, eb1, na1, nb1);
        if (t + 3 < y_hi) load_psi(t + 3, ia1, ib1);
        if (t + 1 < y_hi) load_e(t + 1, ea1, eb1);
    };
    if constexpr (sizeof(T) == 8) {
        // FP64: the steady state (t >= y_lo + 8, t + 3 < y_hi: rows t-4..t+3 in
        // the segment, the emitted rows in the output range) in a loop of its own
        // with every validity known and no conditional memory operation, which
        // lets the compiler fit the pass in 254 VGPRs (2 waves per SIMD; the
        // single loop needed 256 plus AGPR spills, 1 wave): 0.350 -> 0.324 ms per
        // step at 4096^2.
        auto beat_steady = [&](auto tp, int t, Row<T, VEC> &ia, Row<T, VEC> &ib, const Row<T, VEC> &ea,
                               const Row<T, VEC> &eb, T (&na)[VEC], T (&nb)[VEC]) {
            constexpr int TP = decltype(tp)::value;
            if (div0_on) scale_row<T, VEC>(ia, inv0);
            if (div1_on) scale_row<T, VEC>(ib, inv1);
            T ta[VEC], tb[VEC];
            density(ia, ta);
            density(ib, tb);
            Row<T, VEC> oa, ob;
            pair_pipe<TP, IMAG, PH>(qa, ia, ea, nb, oa, true, true, true, true, true, hc0, mV0, tV0, unV0, c0, cab,
                                    mag0);
            pair_pipe<TP, IMAG, PH>(qb, ib, eb, na, ob, true, true, true, true, true, hc1, mV1, tV1, unV1, c1, cab,
                                    mag1);
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                na[k] = ta[k];
                nb[k] = tb[k];
            }
            emit(t - 4, oa, ob);
        };
        for (; t < t_end && t < y_lo + 8; t += 2) pair_general(t);
        for (; t + 3 < y_hi; t += 2) {
            beat_steady(I0(), t, ia0, ib0, ea0, eb0, na0, nb0);
            load_psi(t + 2, ia0, ib0);
            load_e(t, ea0, eb0);
            beat_steady(I1(), t + 1, ia1, ib1, ea1, eb1, na1, nb1);
            load_psi(t + 3, ia1, ib1);
            load_e(t + 1, ea1, eb1);
        }
        for (; t < t_end; t += 2) pair_general(t);
    } else {
        // FP32: one loop (measured faster than the split one in FP32)
        for (; t < t_end; t += 2) {
            beat(I0(), t, ia0, ib0, ea0, eb0, na0, nb0);
            if (t + 2 < y_hi) load_psi(t + 2, ia0, ib0);
            if (t < y_hi) load_e(t, ea0, eb0);
            beat(I1(), t + 1, ia1, ib1, ea1, eb1, na1, nb1);
            if (t + 3 < y_hi) load_psi(t + 3, ia1, ib1);
            if (t + 1 < y_hi) load_e(t + 1, ea1, eb1);
        }
    }

    if (IMAG) {
        // per-component renormalisation (src/cpukernel.cpp:551-563): one partial
        // per wave and component; the wave that finishes last sums them in wave
        // order and publishes both divisors
        double s0 = acc0 + cmp0, s1 = acc1 + cmp1;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o, kWave);
            s1 += __shfl_xor(s1, o, kWave);
        }
        unsigned ticket = 0u;
        if (lane == 0) {
            p.partials[wave] = s0;
            p.partials[p.n_waves + wave] = s1;
            __threadfence();
            ticket = atomicAdd(p.counter, 1u);
        }
        ticket = __builtin_amdgcn_readfirstlane(ticket);
        if (ticket == (unsigned)p.n_waves - 1) {
            __threadfence();
            double a0 = 0.0, e0 = 0.0, a1 = 0.0, e1 = 0.0;
            for (int i = lane; i < p.n_waves; i += kWave) {
                neumaier(a0, e0, __hip_atomic_load(p.partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                neumaier(a1, e1, __hip_atomic_load(p.partials + p.n_waves + i, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT));
            }
            double t0 = a0 + e0, t1 = a1 + e1;
            for (int o = kWave / 2; o > 0; o >>= 1) {
                t0 += __shfl_xor(t0, o, kWave);
                t1 += __shfl_xor(t1, o, kWave);
            }
            if (lane == 0) {
                const double n0 = p.norm[0], n1 = p.norm[1];
                p.div_out[0] = n0 != 0.0 ? sqrt(t0 * p.dxdy / n0) : 1.0;
                p.div_out[1] = n1 != 0.0 ? sqrt(t1 * p.dxdy / n1) : 1.0;
                __hip_atomic_store(p.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// MODE as in pair2_kernel (real time; imaginary time runs -1).
template <typename T, bool IMAG, int VEC, bool PH, int MODE>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 3 : 1) pair_kernel(PairParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(xcd_block<true>(blockIdx.x, gridDim.x) * (blockDim.x / kWave) +
                                                    (threadIdx.x / kWave));
    if (wave >= p.n_waves) return;
    pair_wave<T, IMAG, VEC, PH, MODE>(p, wave, lane);
}

// ---------------------------------------------------------------------------
// Two real-time steps of both components per pass
// ---------------------------------------------------------------------------
// pair_wave's two component pipelines, twice in series with an 8-cell apron (the
// dependency cone of two steps), as step2_wave chains two single-component
// pipelines: the first pair (A) runs step n on the input rows; every pair of rows
// it emits (4 rows behind its input) is Rabi-rotated by the first step's
// constants and enters the second pair (B), which runs step n+1 and emits 8 rows
// behind the input, rotated by the second step's constants.  B's potential
// passes see the densities of B's input rows -- the start-of-step densities of
// step n+1, after the rotation, as in the reference, where the rotation ends a
// step (src/solver.cpp:201-226).  Per site a pass reads psi_a, psi_b once, the E
// rows twice (mostly from the caches) and writes psi_a, psi_b once.  Exactness
// across tiles follows the single-component pass: closed axes, or periodic axes
// with the 8-wide halo refreshed before each pass and E wrapped over it.
#ifndef TSMI_PAIR2_SYNC
#define TSMI_PAIR2_SYNC 4
#endif
// Beats between workgroup barriers in the fast loop of pair2_wave (0: none;
// even): the workgroup's four waves run neighbouring strips, which read the same
// apron columns, and in step they fetch those lines through L2 once.
constexpr int kPair2Sync = TSMI_PAIR2_SYNC;
static_assert(kPair2Sync % 2 == 0, "the fast loop runs two beats per iteration");
constexpr int kPair2Block = 256;     // threads per pair2_kernel workgroup (its launch bound)

template <typename T, int VEC, bool PH, int MODE>
__device__ __forceinline__ void pair2_wave(const PairParams<T> &p, const int wave, const int lane,
                                           const int block_wave0 = -1) {
    constexpr int AP = 2 * kApron;
    constexpr int SW = kWave * VEC - 2 * AP;
    const int strip = wave % p.n_strips;
    const int seg = wave / p.n_strips;
    const int X0 = strip * SW - AP;
    const int xl = X0 + lane * VEC;
    const int y_out_lo = seg * p.seg_rows;
    const int y_out_hi = min(p.H, y_out_lo + p.seg_rows);
    const int y_lo = max(0, y_out_lo - AP);
    const int y_hi = min(p.H, y_out_hi + AP);
    const bool ld_ok = xl >= 0 && xl + VEC <= p.pitch;
    const bool own = lane >= AP / VEC && lane < kWave - AP / VEC;

    const HCoef<T, VEC> hc0 = make_hcoef<T, VEC>(xl, lane, p.W, p.mH[0], p.tH[0], p.unH[0]);
    const HCoef<T, VEC> hc1 = make_hcoef<T, VEC>(xl, lane, p.W, p.mH[1], p.tH[1], p.unH[1]);
    const T mV0 = p.mV[0], tV0 = p.tV[0], unV0 = p.unV[0], c0 = p.c_self[0];
    const T mV1 = p.mV[1], tV1 = p.tV[1], unV1 = p.unV[1], c1 = p.c_self[1];
    const T cab = p.c_ab;
    const double mag0 = p.mag[0], mag1 = p.mag[1];
    const bool rabi = p.rabi != 0;

    const long seg0 = (long)y_lo * p.pitch;
    const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
    const __amdgpu_buffer_rsrc_t ri0r = make_rsrc(p.in_re[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri0i = make_rsrc(p.in_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri1r = make_rsrc(p.in_re[1] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ri1i = make_rsrc(p.in_im[1] + seg0, seg_bytes);
    const T *e0 = PH ? p.ph[0] : p.e_re[0], *e1 = PH ? p.ph[1] : p.e_re[1];
    const __amdgpu_buffer_rsrc_t re0r = make_rsrc(e0 + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re0i = make_rsrc(PH ? e0 + seg0 : p.e_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re1r = make_rsrc(e1 + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t re1i = make_rsrc(PH ? e1 + seg0 : p.e_im[1] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro0r = make_rsrc(p.out_re[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro0i = make_rsrc(p.out_im[0] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro1r = make_rsrc(p.out_re[1] + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t ro1i = make_rsrc(p.out_im[1] + seg0, seg_bytes);
    const unsigned row_bytes = (unsigned)p.pitch * (unsigned)sizeof(T);
    const unsigned v_ld = ld_ok ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    const unsigned v_st = (own && ld_ok) ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;

    auto load_psi = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(ri0r, v_ld, so, a.re);
        buf_load<T, VEC>(ri0i, v_ld, so, a.im);
        buf_load<T, VEC>(ri1r, v_ld, so, b.re);
        buf_load<T, VEC>(ri1i, v_ld, so, b.im);
    };
    // a state row past the segment through the vector offset: the buffer's
    // range check turns it into a load of zeros
    auto load_psi_any = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
        const unsigned off = v_ld + (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(ri0r, off, 0u, a.re);
        buf_load<T, VEC>(ri0i, off, 0u, a.im);
        buf_load<T, VEC>(ri1r, off, 0u, b.re);
        buf_load<T, VEC>(ri1i, off, 0u, b.im);
    };
    auto load_e = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(re0r, v_ld, so, a.re);
        buf_load<T, VEC>(re1r, v_ld, so, b.re);
        if (!PH) {
            buf_load<T, VEC>(re0i, v_ld, so, a.im);
            buf_load<T, VEC>(re1i, v_ld, so, b.im);
        }
    };
    // the second step's rotation, then the stores; a row before the output range
    // gets an offset past the buffer, which drops the store
    auto emit_any = [&](int y, Row<T, VEC> &xa, Row<T, VEC> &xb) {
        if (rabi) rabi_rows<T, VEC>(xa, xb, p.cc2, p.cs_r2, p.cs_i2);
        const unsigned off = y >= y_out_lo && y < y_out_hi ? v_st + (unsigned)(y - y_lo) * row_bytes : kOffNone;
        buf_store<T, VEC>(ro0r, off, 0u, xa.re);
        buf_store<T, VEC>(ro0i, off, 0u, xa.im);
        buf_store<T, VEC>(ro1r, off, 0u, xb.re);
        buf_store<T, VEC>(ro1i, off, 0u, xb.im);
    };

    Pipe4<T, VEC> A0, A1, B0, B1;
    // input rows, E rows of pair A (row t-2) and of pair B (row t-6), and the
    // start-of-step densities of each pair's input rows, in slots by row parity
    Row<T, VEC> ia0, ib0, ia1, ib1, ea0, eb0, ea1, eb1, fa0, fb0, fa1, fb1;
    T na0[VEC], nb0[VEC], na1[VEC], nb1[VEC], ma0[VEC], mb0[VEC], ma1[VEC], mb1[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        A0.p1.re[k] = A0.p1.im[k] = A0.p2.re[k] = A0.p2.im[k] = A0.p3.re[k] = A0.p3.im[k] = T(0);
        A0.p4.re[k] = A0.p4.im[k] = A1.p1.re[k] = A1.p1.im[k] = A1.p2.re[k] = A1.p2.im[k] = T(0);
        A1.p3.re[k] = A1.p3.im[k] = A1.p4.re[k] = A1.p4.im[k] = T(0);
        B0.p1.re[k] = B0.p1.im[k] = B0.p2.re[k] = B0.p2.im[k] = B0.p3.re[k] = B0.p3.im[k] = T(0);
        B0.p4.re[k] = B0.p4.im[k] = B1.p1.re[k] = B1.p1.im[k] = B1.p2.re[k] = B1.p2.im[k] = T(0);
        B1.p3.re[k] = B1.p3.im[k] = B1.p4.re[k] = B1.p4.im[k] = T(0);
        ia0.re[k] = ia0.im[k] = ib0.re[k] = ib0.im[k] = ia1.re[k] = ia1.im[k] = ib1.re[k] = ib1.im[k] = T(0);
        ea0.re[k] = ea0.im[k] = eb0.re[k] = eb0.im[k] = ea1.re[k] = ea1.im[k] = eb1.re[k] = eb1.im[k] = T(0);
        fa0.re[k] = fa0.im[k] = fb0.re[k] = fb0.im[k] = fa1.re[k] = fa1.im[k] = fb1.re[k] = fb1.im[k] = T(0);
        na0[k] = nb0[k] = na1[k] = nb1[k] = ma0[k] = mb0[k] = ma1[k] = mb1[k] = T(0);
    }
    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };

    // beat t with row validities v[k] = valid(t - k): pair A takes input row t
    // and gives row t-4, pair B takes it and gives row t-8 (emitted when wanted)
    auto beat = [&](auto tp, int t, bool v0, bool v1, bool v2, bool v3, bool v4, bool v5, bool v6, bool v7, bool v8,
                    bool want, Row<T, VEC> &ia, Row<T, VEC> &ib, const Row<T, VEC> &ea, const Row<T, VEC> &eb,
                    const Row<T, VEC> &fa, const Row<T, VEC> &fb, T (&na)[VEC], T (&nb)[VEC], T (&ma)[VEC],
                    T (&mb)[VEC]) {
        constexpr int TP = decltype(tp)::value;
        T ta[VEC], tb[VEC];
        density(ia, ta);
        density(ib, tb);
        Row<T, VEC> oa, ob;
        pair_pipe<TP, false, PH>(A0, ia, ea, nb, oa, v0, v1, v2, v3, v4, hc0, mV0, tV0, unV0, c0, cab, mag0);
        pair_pipe<TP, false, PH>(A1, ib, eb, na, ob, v0, v1, v2, v3, v4, hc1, mV1, tV1, unV1, c1, cab, mag1);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            na[k] = ta[k];
            nb[k] = tb[k];
        }
        if (rabi) rabi_rows<T, VEC>(oa, ob, p.cc, p.cs_r, p.cs_i);
        density(oa, ta);
        density(ob, tb);
        Row<T, VEC> xa, xb;
        pair_pipe<TP, false, PH>(B0, oa, fa, mb, xa, v4, v5, v6, v7, v8, hc0, mV0, tV0, unV0, c0, cab, mag0);
        pair_pipe<TP, false, PH>(B1, ob, fb, ma, xb, v4, v5, v6, v7, v8, hc1, mV1, tV1, unV1, c1, cab, mag1);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            ma[k] = ta[k];
            mb[k] = tb[k];
        }
        if (want) emit_any(t - 8, xa, xb);
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    // Every beat in the fast form when the strip lies inside the tile and the
    // segment's 8 apron rows above and below inside the lattice: the zero rows
    // the pipelines start with and the zero rows loaded past the segment meet
    // apron rows only, and a wrong value moves at most one row per vertical
    // pass, eight per pass of two steps, so it never reaches an output row.
    if constexpr (MODE >= 0) {
        // strip inside the tile, apron rows inside the lattice: the fast loop
        auto fast_ok = [&](int w) {
            const int st = w % p.n_strips, sg = w / p.n_strips;
            const int x0 = st * SW - AP, lo = sg * p.seg_rows;
            return w < p.n_waves && x0 >= 0 && x0 + kWave * VEC <= p.W && lo >= AP &&
                   min(p.H, lo + p.seg_rows) + AP <= p.H;
        };
        if (fast_ok(wave)) {
            // all four waves of the workgroup in the fast loop (the same segment
            // height, so the same trip count): a barrier every kPair2Sync beats
            bool sync = false;
            if constexpr (kPair2Sync > 0) {
                sync = block_wave0 >= 0;
#pragma unroll
                for (int i = 0; i < kPair2Block / kWave; ++i) sync = sync && fast_ok(block_wave0 + i);
            }
            const T kH0 = MODE == 0 ? p.tH[0] : p.mH[0], kV0 = MODE == 0 ? p.tV[0] : p.mV[0];
            const T kH1 = MODE == 0 ? p.tH[1] : p.mH[1], kV1 = MODE == 0 ? p.tV[1] : p.mV[1];
            // E rows by row index; a row outside the segment reads zeros
            auto load_e_at = [&](int y, Row<T, VEC> &a, Row<T, VEC> &b) {
                const unsigned off = (y >= y_lo && y < y_hi) ? v_ld + (unsigned)(y - y_lo) * row_bytes : kOffNone;
                buf_load<T, VEC>(re0r, off, 0u, a.re);
                buf_load<T, VEC>(re1r, off, 0u, b.re);
                if (!PH) {
                    buf_load<T, VEC>(re0i, off, 0u, a.im);
                    buf_load<T, VEC>(re1i, off, 0u, b.im);
                }
            };
            auto beat_fast = [&](auto tp, int t, Row<T, VEC> &ia, Row<T, VEC> &ib, const Row<T, VEC> &ea,
                                 const Row<T, VEC> &eb, const Row<T, VEC> &fa, const Row<T, VEC> &fb, T (&na)[VEC],
                                 T (&nb)[VEC], T (&ma)[VEC], T (&mb)[VEC]) {
                constexpr int TP = decltype(tp)::value;
                T ta[VEC], tb[VEC];
                density(ia, ta);
                density(ib, tb);
                Row<T, VEC> oa, ob;
                pair_pipe_fast<TP, MODE, PH>(A0, ia, ea, nb, oa, kH0, kV0, c0, cab, mag0);
                pair_pipe_fast<TP, MODE, PH>(A1, ib, eb, na, ob, kH1, kV1, c1, cab, mag1);
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    na[k] = ta[k];
                    nb[k] = tb[k];
                }
                if (rabi) rabi_rows<T, VEC>(oa, ob, p.cc, p.cs_r, p.cs_i);
                density(oa, ta);
                density(ob, tb);
                Row<T, VEC> xa, xb;
                pair_pipe_fast<TP, MODE, PH>(B0, oa, fa, mb, xa, kH0, kV0, c0, cab, mag0);
                pair_pipe_fast<TP, MODE, PH>(B1, ob, fb, ma, xb, kH1, kV1, c1, cab, mag1);
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    ma[k] = ta[k];
                    mb[k] = tb[k];
                }
                emit_any(t - 8, xa, xb);
            };
            int t = y_lo;
            load_psi(t, ia0, ib0);
            if (t + 1 < y_hi) load_psi(t + 1, ia1, ib1);
            for (int it = 0; t < y_hi + 8; t += 2, ++it) {
                beat_fast(I0(), t, ia0, ib0, ea0, eb0, fa0, fb0, na0, nb0, ma0, mb0);
                load_psi_any(t + 2, ia0, ib0);
                load_e_at(t, ea0, eb0);
                load_e_at(t - 4, fa0, fb0);
                beat_fast(I1(), t + 1, ia1, ib1, ea1, eb1, fa1, fb1, na1, nb1, ma1, mb1);
                load_psi_any(t + 3, ia1, ib1);
                load_e_at(t + 1, ea1, eb1);
                load_e_at(t - 3, fa1, fb1);
                if constexpr (kPair2Sync > 0) {
                    if (sync && it % (kPair2Sync / 2) == kPair2Sync / 2 - 1) __builtin_amdgcn_s_barrier();
                }
            }
            return;
        }
    }
    // two beats per iteration; state row r sits in slot [(r - y_lo) % 2]
    auto pair_general = [&](int t) {
        {
            const int y = t - 8;
            beat(I0(), t, valid(t), valid(t - 1), valid(t - 2), valid(t - 3), valid(t - 4), valid(t - 5),
                 valid(t - 6), valid(t - 7), valid(y), valid(y) && y >= y_out_lo && y < y_out_hi, ia0, ib0, ea0, eb0,
                 fa0, fb0, na0, nb0, ma0, mb0);
        }
        if (t + 2 < y_hi) load_psi(t + 2, ia0, ib0);
        if (t < y_hi) load_e(t, ea0, eb0);
        if (valid(t - 4)) load_e(t - 4, fa0, fb0);
        {
            const int u = t + 1, y = u - 8;
            beat(I1(), u, valid(u), valid(u - 1), valid(u - 2), valid(u - 3), valid(u - 4), valid(u - 5),
                 valid(u - 6), valid(u - 7), valid(y), valid(y) && y >= y_out_lo && y < y_out_hi, ia1, ib1, ea1, eb1,
                 fa1, fb1, na1, nb1, ma1, mb1);
        }
        if (t + 3 < y_hi) load_psi(t + 3, ia1, ib1);
        if (t + 1 < y_hi) load_e(t + 1, ea1, eb1);
        if (valid(t - 3)) load_e(t - 3, fa1, fb1);
    };
    int t = y_lo;
    if (t < y_hi) load_psi(t, ia0, ib0);
    if (t + 1 < y_hi) load_psi(t + 1, ia1, ib1);
    const int t_end = y_hi + 8;
    // the pipelines fill
    for (; t < t_end && t < y_lo + 8; t += 2) pair_general(t);
    // steady state (rows t-8 .. t+1 in the segment, every emitted row before the
    // end of the output range): no conditional memory operation
    for (; t + 1 < y_hi; t += 2) {
        beat(I0(), t, true, true, true, true, true, true, true, true, true, true, ia0, ib0, ea0, eb0, fa0, fb0, na0,
             nb0, ma0, mb0);
        load_psi_any(t + 2, ia0, ib0);
        load_e(t, ea0, eb0);
        load_e(t - 4, fa0, fb0);
        beat(I1(), t + 1, true, true, true, true, true, true, true, true, true, true, ia1, ib1, ea1, eb1, fa1, fb1,
             na1, nb1, ma1, mb1);
        load_psi_any(t + 3, ia1, ib1);
        load_e(t + 1, ea1, eb1);
        load_e(t - 3, fa1, fb1);
    }
    // the pipelines drain
    for (; t < t_end; t += 2) pair_general(t);
}

// MODE: the scaled form both components use on both axes (0 or 1), or -1 when
// they differ (the general beat only).
template <typename T, int VEC, bool PH, int MODE>
__global__ void __launch_bounds__(kPair2Block, 2) pair2_kernel(PairParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave0 = xcd_block<true>(blockIdx.x, gridDim.x) * (blockDim.x / kWave);
    const int wave = __builtin_amdgcn_readfirstlane(wave0 + (threadIdx.x / kWave));
    if (wave >= p.n_waves) return;
    pair2_wave<T, VEC, PH, MODE>(p, wave, lane, wave0);
}

// ---------------------------------------------------------------------------
// Rabi rotation pass (and the pending-divisor scaling)
// ---------------------------------------------------------------------------
template <typename T>
struct RabiParams {
    T *re[2], *im[2];            // both components' tiles, updated in place
    int W, H, pitch;
    double cc, cs_r, cs_i;       // FP64 at every precision (see PairParams)
    int rotate;                  // 0: only divide by the pending divisors
    const double *div_in[2];     // pending divisors (nullptr: 1)
    int normalize;               // the joint renormalisation of src/cpukernel.cpp:610-667
    double *partials;            // [2][gridDim.x]
    unsigned *counter;           // zero between launches
    double *div_out;             // both components' divisor after the joint renormalisation
    double *norm;                // per-component targets, redistributed
    double tot_norm, dxdy;
    int ix0, ix1, iy0, iy1;
};

constexpr int kRabiThreads = 256;

// CPUBlock::rabi_coupling on the whole tiles (src/cpukernel.cpp:581-608,
// src/cpucartesian.cpp:264-290) followed, when `normalize`, by
// CPUBlock::normalization: both tiles divided by sqrt((S_a + S_b) dx dy /
// tot_norm) -- here left pending in div_out -- and the targets redistributed as
// S_a / (S_a + S_b) tot_norm, S_b / (S_a + S_b) tot_norm.
template <typename T, bool IMAG>
__global__ void __launch_bounds__(kRabiThreads) rabi_kernel(RabiParams<T> p) {
    const double d0 = p.div_in[0] ? *p.div_in[0] : 1.0, d1 = p.div_in[1] ? *p.div_in[1] : 1.0;
    const bool div0_on = p.div_in[0] != nullptr, div1_on = p.div_in[1] != nullptr;
    const double cc = p.cc, sr = p.cs_r, si = p.cs_i;
    double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
    const long n = (long)p.W * (long)p.H;
    for (long i = (long)blockIdx.x * kRabiThreads + threadIdx.x; i < n; i += (long)gridDim.x * kRabiThreads) {
        const int y = (int)(i / p.W), x = (int)(i - (long)y * p.W);
        const long k = (long)y * p.pitch + x;
        T ar = p.re[0][k], ai = p.im[0][k], br = p.re[1][k], bi = p.im[1][k];
        if (div0_on) { ar = T(double(ar) / d0); ai = T(double(ai) / d0); }
        if (div1_on) { br = T(double(br) / d1); bi = T(double(bi) / d1); }
        T nar = ar, nai = ai, nbr = br, nbi = bi;
        if (p.rotate) {
            const double xr = ar, xi = ai, yr = br, yi = bi;
            if (IMAG) {   // src/cpucartesian.cpp:278-290
                nar = T(cc * xr + sr * yr - si * yi);
                nai = T(cc * xi + si * yr + sr * yi);
                nbr = T(cc * yr + sr * xr + si * xi);
                nbi = T(cc * yi - si * xr + sr * xi);
            } else {      // src/cpucartesian.cpp:264-276
                nar = T(cc * xr - si * yr - sr * yi);
                nai = T(cc * xi + sr * yr - si * yi);
                nbr = T(cc * yr + si * xr - sr * xi);
                nbi = T(cc * yi + sr * xr + si * xi);
            }
        }
        p.re[0][k] = nar; p.im[0][k] = nai;
        p.re[1][k] = nbr; p.im[1][k] = nbi;
        if (p.normalize && x >= p.ix0 && x < p.ix1 && y >= p.iy0 && y < p.iy1) {
            neumaier(s0, c0, double(nar) * double(nar) + double(nai) * double(nai));
            neumaier(s1, c1, double(nbr) * double(nbr) + double(nbi) * double(nbi));
        }
    }
    if (!p.normalize) return;
    __shared__ double red[2][kRabiThreads];
    __shared__ int last;
    red[0][threadIdx.x] = s0 + c0;
    red[1][threadIdx.x] = s1 + c1;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = 0.0, e0 = 0.0, a1 = 0.0, e1 = 0.0;
        for (int i = 0; i < kRabiThreads; ++i) {
            neumaier(a0, e0, red[0][i]);
            neumaier(a1, e1, red[1][i]);
        }
        p.partials[blockIdx.x] = a0 + e0;
        p.partials[gridDim.x + blockIdx.x] = a1 + e1;
        __threadfence();
        last = atomicAdd(p.counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double a0 = 0.0, e0 = 0.0, a1 = 0.0, e1 = 0.0;
        for (unsigned i = 0; i < gridDim.x; ++i) {
            neumaier(a0, e0, __hip_atomic_load(p.partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            neumaier(a1, e1, __hip_atomic_load(p.partials + gridDim.x + i, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT));
        }
        const double sa = a0 + e0, sb = a1 + e1;
        const double f = sqrt((sa + sb) * p.dxdy / p.tot_norm);
        p.div_out[0] = f;
        p.div_out[1] = f;
        p.norm[0] = sa / (sa + sb) * p.tot_norm;
        p.norm[1] = sb / (sa + sb) * p.tot_norm;
        __hip_atomic_store(p.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Cross sums of the two-component energy (src/solver.cpp:340-354) over the
// inner region: sum |a|^2, sum |b|^2, sum |a|^4 |b|^4, sum |a|^2 |b|^2 2 Re(conj(a) b omega).
template <typename T>
__global__ void __launch_bounds__(kRabiThreads) pair_cross_kernel(const T *ar_, const T *ai_, const T *br_,
                                                                   const T *bi_, int pitch, int ix0, int ix1,
                                                                   int iy0, int iy1, double om_r, double om_i,
                                                                   double *partials) {
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int w = ix1 - ix0;
    const long n = (long)w * (long)(iy1 - iy0);
    for (long i = (long)blockIdx.x * kRabiThreads + threadIdx.x; i < n; i += (long)gridDim.x * kRabiThreads) {
        const int y = iy0 + (int)(i / w), x = ix0 + (int)(i % w);
        const long k = (long)y * pitch + x;
        const double ar = ar_[k], ai = ai_[k], br = br_[k], bi = bi_[k];
        const double na = ar * ar + ai * ai, nb = br * br + bi * bi;
        // conj(a) b = (ar br + ai bi) + i (ar bi - ai br)
        const double xr = ar * br + ai * bi, xi = ar * bi - ai * br;
        neumaier(s[0], c[0], na);
        neumaier(s[1], c[1], nb);
        neumaier(s[2], c[2], na * na * nb * nb);
        neumaier(s[3], c[3], na * nb * 2.0 * (xr * om_r - xi * om_i));
    }
    __shared__ double red[4][kRabiThreads];
#pragma unroll
    for (int j = 0; j < 4; ++j) red[j][threadIdx.x] = s[j] + c[j];
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0.0, e = 0.0;
        for (int t = 0; t < kRabiThreads; ++t) neumaier(a, e, red[threadIdx.x][t]);
        partials[(long)blockIdx.x * 4 + threadIdx.x] = a + e;
    }
}

}  // namespace tsmi
