// adjoint_kernel.hpp -- one reverse step of the real-time evolution, with the
// gradient of the loss with respect to the potential.
//
// No reference counterpart: the reference evolves forward only.  A real-time
// step is psi_{n+1} = A N(A' psi_n), A' the first four kinetic passes
// V0 H0 V1 H1, A the last four, N the potential stage
// N(chi) = E chi e^(-i phi(|chi|^2)), phi(s) = c s + c_LHY s^(3/2),
// E = exp(-i dt V).  Every kinetic pass is unitary and symmetric, so its inverse
// and its adjoint are the pass with b -> -b; |N(chi)| = |chi|, so N^-1(y) =
// conj(E) e^(i phi(|y|^2)) y.  The step at -dt (b -> -b, conj(E), -phi) is
// therefore the inverse step.  With the adjoint state
// G = dL/dRe psi + i dL/dIm psi and the two states at the middle of the reverse
// step, mu = A^H G_{n+1} and y = A^-1 psi_{n+1} (the forward's post-potential
// state), at every cell
//     dL/dV += dt Im(conj(mu) y),
//     N^T mu = conj(E) e^(i phi) mu + 2 phi'(s) Im(conj(mu) y) x,
// x = N^-1(y), s = |y|^2, phi'(s) = c + 1.5 c_LHY s^(1/2); with c = c_LHY = 0
// this is conj(E) mu.  The kernel runs the step at -dt on two states at once --
// G, and psi_{n+1}, which comes back as psi_n -- and at the potential stage
// adds dt Im(conj(mu) y) into an FP64 gradient plane.  The host re-anchors psi
// on the checkpoints of the forward sweep, so reversal round-off never
// accumulates over more than one checkpoint interval.
//
// The layout, strip/segment decomposition and pipeline are those of
// step_kernel; the scaled-frame factors (m, t, un) are the host's values for
// the step at -dt.  In that frame the two mid-step states are y/conj(kappa)
// and mu/conj(kappa), kappa = (gH gV)^2, so gscale = dt |kappa|^2 restores the
// gradient, the couplings are the forward's mid-step ones (c |kappa|^2,
// c_LHY |kappa|^3) and the E planes carry kappa^2.
#pragma once
#include "step_kernel.hpp"

namespace tsmi {

template <typename T>
struct AdjParams {
    const T *s_in_re, *s_in_im;      // psi_{n+1}
    T *s_out_re, *s_out_im;          // psi_n
    const T *g_in_re, *g_in_im;      // G_{n+1}
    T *g_out_re, *g_out_im;          // G_n
    const T *e_re, *e_im;            // the forward E planes (E times the forward frame factor)
    double *grad;                    // dL/dV accumulator, FP64, same pitch
    int W, H, pitch;
    int n_strips, n_waves, seg_rows;
    T mH, tH, mV, tV, unH, unV;      // frame factors of the step at -dt
    T c_nl, c_lhy;                   // couplings in the mid-step frame
    double gscale;
};

template <typename T, int VEC>
struct Pair2 {
    Row<T, VEC> s, g;
};

template <typename T, int VEC>
__device__ __forceinline__ void conj_e_mul(Row<T, VEC> &r, const Row<T, VEC> &e) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T pr = r.re[k], pi = r.im[k];
        r.re[k] = e.re[k] * pr + e.im[k] * pi;
        r.im[k] = e.re[k] * pi - e.im[k] * pr;
    }
}

// N^-1 on the state row s (y -> x) and N^T on the adjoint row g (mu), in place.
template <typename T, int VEC>
__device__ __forceinline__ void nl_reverse(Row<T, VEC> &s, Row<T, VEC> &g, const Row<T, VEC> &e, T c, T cl) {
    T ang[VEC], f[VEC], sn[VEC], cs[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T n2 = s.re[k] * s.re[k] + s.im[k] * s.im[k];
        const T r = sqrt_t(n2);
        ang[k] = c * n2 + cl * (n2 * r);
        f[k] = T(2) * (c + T(1.5) * cl * r) * (g.re[k] * s.im[k] - g.im[k] * s.re[k]);
    }
    sincos_row<T, VEC>(ang, sn, cs);
    conj_e_mul(s, e);
    conj_e_mul(g, e);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        const T sr = cs[k] * s.re[k] - sn[k] * s.im[k], si = sn[k] * s.re[k] + cs[k] * s.im[k];
        const T gr = cs[k] * g.re[k] - sn[k] * g.im[k], gi = sn[k] * g.re[k] + cs[k] * g.im[k];
        s.re[k] = sr;
        s.im[k] = si;
        g.re[k] = fma(f[k], sr, gr);
        g.im[k] = fma(f[k], si, gi);
    }
}

template <typename T, int VEC>
__device__ __forceinline__ void buf_load_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double (&v)[VEC]) {
    static_assert(VEC % 2 == 0, "gradient rows are loaded in 16-byte pieces");
#pragma unroll
    for (int i = 0; i < VEC / 2; ++i) {
        const u32x4 w = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * i, soff, 0));
        typedef double __attribute__((ext_vector_type(2))) D2;
        const D2 x = __builtin_bit_cast(D2, w);
        v[2 * i] = x[0];
        v[2 * i + 1] = x[1];
    }
}

template <typename T, int VEC>
__device__ __forceinline__ void buf_store_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const double (&v)[VEC]) {
#pragma unroll
    for (int i = 0; i < VEC / 2; ++i) {
        typedef double __attribute__((ext_vector_type(2))) D2;
        D2 x;
        x[0] = v[2 * i];
        x[1] = v[2 * i + 1];
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, voff + soff + 16 * i, 0, 0);
    }
}

template <typename T, int VEC, bool NL>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave));
    if (wave >= p.n_waves) return;
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

    const HCoef<T, VEC> hc = make_hcoef<T, VEC>(xl, lane, p.W, p.mH, p.tH, p.unH);
    const T mV = p.mV, tV = p.tV, unV = p.unV;

    const long seg0 = (long)y_lo * p.pitch;
    const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
    const unsigned seg_bytes_d = (unsigned)((long)(y_hi - y_lo) * p.pitch * 8L);
    const __amdgpu_buffer_rsrc_t r_si_re = make_rsrc(p.s_in_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_si_im = make_rsrc(p.s_in_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_gi_re = make_rsrc(p.g_in_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_gi_im = make_rsrc(p.g_in_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_so_re = make_rsrc(p.s_out_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_so_im = make_rsrc(p.s_out_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_go_re = make_rsrc(p.g_out_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_go_im = make_rsrc(p.g_out_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_e_re = make_rsrc(p.e_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_e_im = make_rsrc(p.e_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_grad = make_rsrc(p.grad + seg0, seg_bytes_d);
    const unsigned row_bytes = (unsigned)p.pitch * (unsigned)sizeof(T);
    const unsigned row_bytes_d = (unsigned)p.pitch * 8u;
    const unsigned v_ld = ld_ok ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    const unsigned v_st = (own && ld_ok) ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    const unsigned v_gd = (own && ld_ok) ? (unsigned)xl * 8u : kOffNone;

    auto load_in = [&](int y, Pair2<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_si_re, v_ld, so, r.s.re);
        buf_load<T, VEC>(r_si_im, v_ld, so, r.s.im);
        buf_load<T, VEC>(r_gi_re, v_ld, so, r.g.re);
        buf_load<T, VEC>(r_gi_im, v_ld, so, r.g.im);
    };
    auto load_e = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_e_re, v_ld, so, r.re);
        buf_load<T, VEC>(r_e_im, v_ld, so, r.im);
    };
    auto load_grad = [&](int y, double (&gr)[VEC]) {
        buf_load_f64<T, VEC>(r_grad, v_gd, (unsigned)(y - y_lo) * row_bytes_d, gr);
    };

    Pair2<T, VEC> pend1, pend2, pend3, pend4, inA, inB;
    Row<T, VEC> eA, eB;
    double gA[VEC], gB[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        pend1.s.re[k] = pend1.s.im[k] = pend1.g.re[k] = pend1.g.im[k] = T(0);
        pend2.s.re[k] = pend2.s.im[k] = pend2.g.re[k] = pend2.g.im[k] = T(0);
        pend3.s.re[k] = pend3.s.im[k] = pend3.g.re[k] = pend3.g.im[k] = T(0);
        pend4.s.re[k] = pend4.s.im[k] = pend4.g.re[k] = pend4.g.im[k] = T(0);
        inA.s.re[k] = inA.s.im[k] = inA.g.re[k] = inA.g.im[k] = T(0);
        inB.s.re[k] = inB.s.im[k] = inB.g.re[k] = inB.g.im[k] = T(0);
        eA.re[k] = eA.im[k] = eB.re[k] = eB.im[k] = T(0);
        gA[k] = gB[k] = 0.0;
    }

    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };

    auto beat = [&](auto tp, int t, Pair2<T, VEC> &in, const Row<T, VEC> &e, double (&gr)[VEC]) {
        constexpr int TP = decltype(tp)::value;
        const bool v0 = valid(t), v1 = valid(t - 1), v2 = valid(t - 2), v3 = valid(t - 3), v4 = valid(t - 4);
        const T a1 = (v0 && v1) ? mV : unV, b1 = (v0 && v1) ? tV : T(0);
        const T a2 = (v1 && v2) ? mV : unV, b2 = (v1 && v2) ? tV : T(0);
        const T a3 = (v2 && v3) ? mV : unV, b3 = (v2 && v3) ? tV : T(0);
        const T a4 = (v3 && v4) ? mV : unV, b4 = (v3 && v4) ? tV : T(0);
        // stage 1: V0, then H0 on row t-1
        vpair<(TP + 1) & 1, false>(pend1.s, in.s, a1, b1);
        vpair<(TP + 1) & 1, false>(pend1.g, in.g, a1, b1);
        Pair2<T, VEC> o1 = pend1;
        pend1 = in;
        hpass<(TP + 1) & 1, false>(o1.s, hc);
        hpass<(TP + 1) & 1, false>(o1.g, hc);
        // stage 2: V1, H1; the gradient; conj(E); H1 on row t-2
        vpair<(TP + 1) & 1, false>(pend2.s, o1.s, a2, b2);
        vpair<(TP + 1) & 1, false>(pend2.g, o1.g, a2, b2);
        Pair2<T, VEC> o2 = pend2;
        pend2 = o1;
        hpass<(TP + 1) & 1, false>(o2.s, hc);
        hpass<(TP + 1) & 1, false>(o2.g, hc);
        const int y2 = t - 2;
        if (v2 && y2 >= y_out_lo && y2 < y_out_hi) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                gr[k] = fma(p.gscale, double(o2.g.re[k]) * double(o2.s.im[k]) - double(o2.g.im[k]) * double(o2.s.re[k]),
                            gr[k]);
            buf_store_f64<T, VEC>(r_grad, v_gd, (unsigned)(y2 - y_lo) * row_bytes_d, gr);
        }
        if constexpr (NL) {
            nl_reverse<T, VEC>(o2.s, o2.g, e, p.c_nl, p.c_lhy);
        } else {
            conj_e_mul(o2.s, e);
            conj_e_mul(o2.g, e);
        }
        hpass<(TP + 1) & 1, false>(o2.s, hc);
        hpass<(TP + 1) & 1, false>(o2.g, hc);
        // stage 3: V1, then H0 on row t-3
        vpair<TP & 1, false>(pend3.s, o2.s, a3, b3);
        vpair<TP & 1, false>(pend3.g, o2.g, a3, b3);
        Pair2<T, VEC> o3 = pend3;
        pend3 = o2;
        hpass<(TP + 1) & 1, false>(o3.s, hc);
        hpass<(TP + 1) & 1, false>(o3.g, hc);
        // stage 4: V0; row t-4 leaves
        vpair<TP & 1, false>(pend4.s, o3.s, a4, b4);
        vpair<TP & 1, false>(pend4.g, o3.g, a4, b4);
        const Pair2<T, VEC> o4 = pend4;
        pend4 = o3;
        const int y = t - 4;
        if (v4 && y >= y_out_lo && y < y_out_hi) {
            const unsigned so = (unsigned)(y - y_lo) * row_bytes;
            buf_store<T, VEC>(r_so_re, v_st, so, o4.s.re);
            buf_store<T, VEC>(r_so_im, v_st, so, o4.s.im);
            buf_store<T, VEC>(r_go_re, v_st, so, o4.g.re);
            buf_store<T, VEC>(r_go_im, v_st, so, o4.g.im);
        }
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    int t = y_lo;
    if (t < y_hi) load_in(t, inA);
    if (t + 1 < y_hi) load_in(t + 1, inB);
    const int t_end = y_hi + 4;
    for (; t < t_end; t += 2) {
        beat(I0(), t, inA, eA, gA);
        if (t + 2 < y_hi) load_in(t + 2, inA);
        if (t < y_hi) { load_e(t, eA); load_grad(t, gA); }
        beat(I1(), t + 1, inB, eB, gB);
        if (t + 3 < y_hi) load_in(t + 3, inB);
        if (t + 1 < y_hi) { load_e(t + 1, eB); load_grad(t + 1, gB); }
    }
}

}  // namespace tsmi
