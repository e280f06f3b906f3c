// fused_kernel.hpp -- two real-time steps per pass over HBM (temporal blocking).
//
// The step kernel's strip/segment decomposition with an 8-cell apron on each
// side (the dependency cone of two steps) and two register pipelines in
// series: pipeline A runs the nine sub-passes of step n on the input rows, and
// every row it emits (4 rows behind its input) is the input row of pipeline B,
// which runs step n+1 and emits 8 rows behind the input.  Each state row is read
// once and written once for two steps.  FP32 segments strictly inside the tile
// (all but the first and last strip and segment) run step2_stream_f32, which
// keeps every row in fixed registers and reads each E row once; the other
// segments run step2_wave, whose pipelines read each E row twice, 4 rows apart.
//
// Two steps on a tile with closed edges equal two single steps exactly when no
// halo exchange is due between them: on closed axes always; on periodic axes
// the host keeps an 8-wide halo, refreshes it before each pass (so the 4-wide
// reference halo the second step reads holds the wrapped values the exchange
// would have written) and gives the pass the fully wrapped E table.  Real time
// only: imaginary time renormalises between steps.
#pragma once
#include "step_kernel.hpp"

#include <utility>

namespace tsmi {

template <typename T, int VEC>
struct Pipe {
    Row<T, VEC> p1, p2, p3, p4;
};

// One beat of a real-time step pipeline (the generic form of step_wave's beat):
// input row t (in), E row t-2 (e), row validity v[k] = valid(t-k); returns row
// t-4 in `out`.
template <int TP, bool LHY, bool PH, typename T, int VEC>
__device__ __forceinline__ void pipe_beat(Pipe<T, VEC> &q, Row<T, VEC> in, const Row<T, VEC> &e, Row<T, VEC> &out,
                                          bool v0, bool v1, bool v2, bool v3, bool v4, const HCoef<T, VEC> &hc,
                                          T mV, T tV, T unV, T c, T cl, double mag) {
    const T a1 = (v0 && v1) ? mV : unV, b1 = (v0 && v1) ? tV : T(0);
    const T a2 = (v1 && v2) ? mV : unV, b2 = (v1 && v2) ? tV : T(0);
    const T a3 = (v2 && v3) ? mV : unV, b3 = (v2 && v3) ? tV : T(0);
    const T a4 = (v3 && v4) ? mV : unV, b4 = (v3 && v4) ? tV : T(0);
    vpair<(TP + 1) & 1, false>(q.p1, in, a1, b1);
    Row<T, VEC> o1 = q.p1;
    q.p1 = in;
    hpass<(TP + 1) & 1, false>(o1, hc);
    vpair<(TP + 1) & 1, false>(q.p2, o1, a2, b2);
    Row<T, VEC> o2 = q.p2;
    q.p2 = o1;
    hpass<(TP + 1) & 1, false>(o2, hc);
    ppass<false, LHY, PH>(o2, e, c, cl, mag);
    hpass<(TP + 1) & 1, false>(o2, hc);
    vpair<TP & 1, false>(q.p3, o2, a3, b3);
    Row<T, VEC> o3 = q.p3;
    q.p3 = o2;
    hpass<(TP + 1) & 1, false>(o3, hc);
    vpair<TP & 1, false>(q.p4, o3, a4, b4);
    out = q.p4;
    q.p4 = o3;
}

// The same beat when the strip is interior and all five rows are valid.
template <int TP, bool LHY, int MODE, bool PH, typename T, int VEC>
__device__ __forceinline__ void pipe_beat_fast(Pipe<T, VEC> &q, Row<T, VEC> in, const Row<T, VEC> &e,
                                               Row<T, VEC> &out, T kH, T kV, T c, T cl, double mag) {
    constexpr int M = MODE < 0 ? 0 : MODE;
    vpair_fast<(TP + 1) & 1, false, M>(q.p1, in, kV);
    Row<T, VEC> o1 = q.p1;
    q.p1 = in;
    hpass_fast<(TP + 1) & 1, false, M>(o1, kH);
    vpair_fast<(TP + 1) & 1, false, M>(q.p2, o1, kV);
    Row<T, VEC> o2 = q.p2;
    q.p2 = o1;
    hpass_fast<(TP + 1) & 1, false, M>(o2, kH);
    ppass<false, LHY, PH>(o2, e, c, cl, mag);
    hpass_fast<(TP + 1) & 1, false, M>(o2, kH);
    vpair_fast<TP & 1, false, M>(q.p3, o2, kV);
    Row<T, VEC> o3 = q.p3;
    q.p3 = o2;
    hpass_fast<(TP + 1) & 1, false, M>(o3, kH);
    vpair_fast<TP & 1, false, M>(q.p4, o3, kV);
    out = q.p4;
    q.p4 = o3;
}

template <typename F, int... K>
__device__ __forceinline__ void unroll_beats(F &&f, std::integer_sequence<int, K...>) {
    (f(std::integral_constant<int, K>()), ...);
}

// The FP32 stream of an interior segment of an interior strip (phase plane, one
// scaled form on both axes): the segment [y_lo, y_hi) lies strictly inside the
// tile, so its first and last 8 rows are apron whatever they hold.  Every beat
// is the fast beat: rows before the segment start as zeros in the pipelines and
// rows past it load as zeros (the buffer range check), which changes only apron
// rows (the dependency cone of two steps is 8 rows).  Rows stay in place: with
// state rows loaded D beats ahead of pipeline A, row r lives in slot r mod
// (D + 8) from its load until it leaves pipeline B (8 beats after entering A),
// and the phase row r in slot r mod (D + 8) from its load until pipeline B's use
// 4 beats after pipeline A's, so pipeline B reads no phase row from memory.
// The loop body is one period of D + 8 beats, so every row keeps its registers
// across the back edge (no copies between pipeline stages).  Beat j (relative
// to y_lo): pipeline A takes row j, uses phase row j-2 and completes row j-4,
// which pipeline B takes (phase row j-6), completing row j-8; then rows j+D
// (state) and j+D-2 (phase) are loaded and the row j-8 is stored if it is an
// output row.
#ifndef TSMI_STREAM_AHEAD
#define TSMI_STREAM_AHEAD 4
#endif
#ifndef TSMI_STREAM
#define TSMI_STREAM 1
#endif
constexpr int kStreamAhead = TSMI_STREAM_AHEAD;
// Beats between workgroup barriers in the stream (0: none; a multiple of the
// period D + 8 beyond one period).  The four waves of a workgroup run four
// neighbouring strips, which read the same apron columns of the same rows: kept
// in step, they fetch those lines through L2 once (8192^2 FP32, per launch:
// FETCH_SIZE 553581 -> 530514 KiB, 317.2 -> 308.6 us; a barrier every 1, 2, 3,
// 4, 6, 24, 36, 48 beats measured 314.1, 310.9, 310.2, 309.7, 310.5, 309.4,
// 310.3, 311.6 us; 8- and 16-wave workgroups 309.6 and 311.9 us).
#ifndef TSMI_STREAM_SYNC
#define TSMI_STREAM_SYNC 12
#endif
constexpr int kStreamSync = TSMI_STREAM_SYNC;
constexpr int kStep2Block = 256;     // threads per step2_kernel workgroup (its launch bound)

template <int TP, bool LHY, int MODE>
__device__ __forceinline__ void stream_pipe(Row<float, 2> &r0, Row<float, 2> &r1, Row<float, 2> &r2,
                                            Row<float, 2> &r3, Row<float, 2> &r4, const Row<float, 2> &e, float kH,
                                            float kV, float c, float cl, double mag) {
    constexpr int M = MODE;
    vpair_fast<(TP + 1) & 1, false, M>(r1, r0, kV);
    hpass_fast<(TP + 1) & 1, false, M>(r1, kH);
    vpair_fast<(TP + 1) & 1, false, M>(r2, r1, kV);
    hpass_fast<(TP + 1) & 1, false, M>(r2, kH);
    ppass<false, LHY, true>(r2, e, c, cl, mag);
    hpass_fast<(TP + 1) & 1, false, M>(r2, kH);
    vpair_fast<TP & 1, false, M>(r3, r2, kV);
    hpass_fast<(TP + 1) & 1, false, M>(r3, kH);
    vpair_fast<TP & 1, false, M>(r4, r3, kV);
}

template <bool LHY, int MODE>
__device__ __forceinline__ void step2_stream_f32(__amdgpu_buffer_rsrc_t r_in_re, __amdgpu_buffer_rsrc_t r_in_im,
                                                 __amdgpu_buffer_rsrc_t r_ph, __amdgpu_buffer_rsrc_t r_out_re,
                                                 __amdgpu_buffer_rsrc_t r_out_im, unsigned v_ld, unsigned v_st,
                                                 unsigned row_bytes, int rows, int out_lo, int out_hi, float kH,
                                                 float kV, float c, float cl, double mag, bool sync) {
    constexpr int D = kStreamAhead;
    constexpr int S = D + 8;
    static_assert(D >= 2 && S % 2 == 0, "pair parities need an even period");
    Row<float, 2> s[S];
    Row<float, 2> e[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
        s[i].re[0] = s[i].re[1] = s[i].im[0] = s[i].im[1] = 0.f;
        e[i].re[0] = e[i].re[1] = e[i].im[0] = e[i].im[1] = 0.f;
    }
    auto row_off = [&](int r) { return (unsigned)r * row_bytes; };
    auto load_state = [&](int r, Row<float, 2> &d) {
        const unsigned off = v_ld + row_off(r);
        buf_load<float, 2>(r_in_re, off, 0u, d.re);
        buf_load<float, 2>(r_in_im, off, 0u, d.im);
    };
    auto load_phase = [&](int r, Row<float, 2> &d) { buf_load<float, 2>(r_ph, v_ld + row_off(r), 0u, d.re); };
#pragma unroll
    for (int r = 0; r < D; ++r) load_state(r, s[r]);
#pragma unroll
    for (int r = 0; r < D - 2; ++r) load_phase(r, e[r]);
    const int beats = (rows + S - 1) / S * S;
    for (int j0 = 0; j0 < beats; j0 += S) {
        unroll_beats(
            [&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int TP = k & 1;
                const int j = j0 + k;
                stream_pipe<TP, LHY, MODE>(s[k % S], s[(k + S - 1) % S], s[(k + S - 2) % S], s[(k + S - 3) % S],
                                           s[(k + S - 4) % S], e[(k + S - 2) % S], kH, kV, c, cl, mag);
                stream_pipe<TP, LHY, MODE>(s[(k + S - 4) % S], s[(k + S - 5) % S], s[(k + S - 6) % S],
                                           s[(k + S - 7) % S], s[(k + S - 8) % S], e[(k + S - 6) % S], kH, kV, c,
                                           cl, mag);
                // row j-8 leaves pipeline B; its slot takes row j+D
                Row<float, 2> &out = s[(k + S - 8) % S];
                const int y = j - 8;
                const unsigned off = (y >= out_lo && y < out_hi) ? v_st + row_off(y) : kOffNone;
                buf_store<float, 2>(r_out_re, off, 0u, out.re);
                buf_store<float, 2>(r_out_im, off, 0u, out.im);
                load_state(j + D, out);
                load_phase(j + D - 2, e[(k + D - 2) % S]);
                if constexpr (kStreamSync > 0 && kStreamSync < S && k % kStreamSync == kStreamSync - 1) {
                    if (sync) __builtin_amdgcn_s_barrier();
                } else if constexpr (kStreamSync >= S && k == S - 1) {
                    static_assert(kStreamSync % S == 0, "a barrier interval beyond one period is whole periods");
                    if (sync && (j0 / S) % (kStreamSync / S) == kStreamSync / S - 1) __builtin_amdgcn_s_barrier();
                }
            },
            std::make_integer_sequence<int, S>());
    }
}

// Two real-time steps of one wavefront's strip segment.  p.seg_rows, p.n_strips
// and p.n_waves describe the fused decomposition (strips of kWave*VEC - 16
// output columns).
template <typename T, int VEC, bool LHY, int MODE, bool PH = false, int SC = 0>
__device__ __forceinline__ void step2_wave(const StepParams<T> &p, const int wave, const int lane) {
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

    const HCoef<T, VEC> hc = make_hcoef<T, VEC>(xl, lane, p.W, p.mH, p.tH, p.unH);
    const T mV = p.mV, tV = p.tV, unV = p.unV;
    const T kH = MODE == 0 ? p.tH : p.mH, kV = MODE == 0 ? p.tV : p.mV;
    const T c = p.c_nl, cl = p.c_lhy;
    const bool interior = MODE >= 0 && X0 >= 0 && X0 + kWave * VEC <= p.W;

    const long seg0 = (long)y_lo * p.pitch;
    const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
    const __amdgpu_buffer_rsrc_t r_in_re = make_rsrc(p.in_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_in_im = make_rsrc(p.in_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_e_re = make_rsrc(PH ? p.ph + seg0 : p.e_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_e_im = make_rsrc(PH ? p.ph + seg0 : p.e_im + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_out_re = make_rsrc(p.out_re + seg0, seg_bytes);
    const __amdgpu_buffer_rsrc_t r_out_im = make_rsrc(p.out_im + seg0, seg_bytes);
    const unsigned row_bytes = (unsigned)p.pitch * (unsigned)sizeof(T);
    const unsigned v_ld = ld_ok ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;
    const unsigned v_st = (own && ld_ok) ? (unsigned)xl * (unsigned)sizeof(T) : kOffNone;

    auto load_row = [&](__amdgpu_buffer_rsrc_t rr, __amdgpu_buffer_rsrc_t ri, int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(rr, v_ld, so, r.re);
        buf_load<T, VEC>(ri, v_ld, so, r.im);
    };
    // E row (both planes), or its phase row (one plane) in PH form
    auto load_e = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_e_re, v_ld, so, r.re);
        if (!PH) buf_load<T, VEC>(r_e_im, v_ld, so, r.im);
    };
    auto valid = [&](int y) { return y >= y_lo && y < y_hi; };
    const double mag = p.mag;

    Pipe<T, VEC> A, B;
    Row<T, VEC> inA, inB, eA, eB, fA, fB;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        A.p1.re[k] = A.p1.im[k] = A.p2.re[k] = A.p2.im[k] = T(0);
        A.p3.re[k] = A.p3.im[k] = A.p4.re[k] = A.p4.im[k] = T(0);
        B.p1.re[k] = B.p1.im[k] = B.p2.re[k] = B.p2.im[k] = T(0);
        B.p3.re[k] = B.p3.im[k] = B.p4.re[k] = B.p4.im[k] = T(0);
        inA.re[k] = inA.im[k] = inB.re[k] = inB.im[k] = T(0);
        eA.re[k] = eA.im[k] = eB.re[k] = eB.im[k] = T(0);
        fA.re[k] = fA.im[k] = fB.re[k] = fB.im[k] = T(0);
    }
    auto emit = [&](int y, const Row<T, VEC> &o) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_store<T, VEC, SC>(r_out_re, v_st, so, o.re);
        buf_store<T, VEC, SC>(r_out_im, v_st, so, o.im);
    };

    // beat t: pipeline A takes input row t (E row t-2) and gives row t-4, the
    // input of pipeline B (E row t-6), which gives row t-8
    auto beat = [&](auto tp, int t, const Row<T, VEC> &in, const Row<T, VEC> &ea, const Row<T, VEC> &eb) {
        constexpr int TP = decltype(tp)::value;
        Row<T, VEC> a, b;
        pipe_beat<TP, LHY, PH>(A, in, ea, a, valid(t), valid(t - 1), valid(t - 2), valid(t - 3), valid(t - 4), hc, mV,
                               tV, unV, c, cl, mag);
        pipe_beat<TP, LHY, PH>(B, a, eb, b, valid(t - 4), valid(t - 5), valid(t - 6), valid(t - 7), valid(t - 8), hc,
                               mV, tV, unV, c, cl, mag);
        const int y = t - 8;
        if (valid(y) && y >= y_out_lo && y < y_out_hi) emit(y, b);
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    // Two beats per iteration; state row r sits in slot in[(r - y_lo) % 2].
    auto pair_general = [&](int t) {
        beat(I0(), t, inA, eA, fA);
        if (t + 2 < y_hi) load_row(r_in_re, r_in_im, t + 2, inA);
        if (t < y_hi) load_e(t, eA);
        if (valid(t - 4)) load_e(t - 4, fA);
        beat(I1(), t + 1, inB, eB, fB);
        if (t + 3 < y_hi) load_row(r_in_re, r_in_im, t + 3, inB);
        if (t + 1 < y_hi) load_e(t + 1, eB);
        if (valid(t - 3)) load_e(t - 3, fB);
    };
    // The steady state of an interior strip (t >= y_lo + 8, t + 1 < y_hi).  No
    // memory operation is conditional: a state row past the segment is
    // addressed through the vector offset, which the buffer's range check turns
    // into a load of zeros, and a row that leaves the second pipeline before
    // the output range (the first four beats) gets an offset past the buffer,
    // which drops the store.
    auto load_row_any = [&](int y, Row<T, VEC> &r) {
        const unsigned so = (unsigned)(y - y_lo) * row_bytes;
        buf_load<T, VEC>(r_in_re, v_ld + so, 0u, r.re);
        buf_load<T, VEC>(r_in_im, v_ld + so, 0u, r.im);
    };
    auto emit_any = [&](int y, const Row<T, VEC> &o) {
        const unsigned off = y >= y_out_lo ? v_st + (unsigned)(y - y_lo) * row_bytes : kOffNone;
        buf_store<T, VEC, SC>(r_out_re, off, 0u, o.re);
        buf_store<T, VEC, SC>(r_out_im, off, 0u, o.im);
    };
    auto pair_fast = [&](int t) {
        Row<T, VEC> a, b;
        pipe_beat_fast<0, LHY, MODE, PH>(A, inA, eA, a, kH, kV, c, cl, mag);
        pipe_beat_fast<0, LHY, MODE, PH>(B, a, fA, b, kH, kV, c, cl, mag);
        emit_any(t - 8, b);
        load_row_any(t + 2, inA);
        load_e(t, eA);
        load_e(t - 4, fA);
        pipe_beat_fast<1, LHY, MODE, PH>(A, inB, eB, a, kH, kV, c, cl, mag);
        pipe_beat_fast<1, LHY, MODE, PH>(B, a, fB, b, kH, kV, c, cl, mag);
        emit_any(t - 7, b);
        load_row_any(t + 3, inB);
        load_e(t + 1, eB);
        load_e(t - 3, fB);
    };
    int t = y_lo;
    if (t < y_hi) load_row(r_in_re, r_in_im, t, inA);
    if (t + 1 < y_hi) load_row(r_in_re, r_in_im, t + 1, inB);
    const int t_end = y_hi + 8;
    if constexpr (sizeof(T) == 4) {
        // FP32 (measured fastest at 8192^2, 4 waves per SIMD): one loop that
        // picks the fast or the general pair of beats per iteration.
        for (; t < t_end; t += 2) {
            if (interior && t >= y_lo + 8 && t + 1 < y_hi) {
                Row<T, VEC> a, b;
                pipe_beat_fast<0, LHY, MODE, PH>(A, inA, eA, a, kH, kV, c, cl, mag);
                pipe_beat_fast<0, LHY, MODE, PH>(B, a, fA, b, kH, kV, c, cl, mag);
                if (t - 8 >= y_out_lo) emit(t - 8, b);
                if (t + 2 < y_hi) load_row(r_in_re, r_in_im, t + 2, inA);
                load_e(t, eA);
                load_e(t - 4, fA);
                pipe_beat_fast<1, LHY, MODE, PH>(A, inB, eB, a, kH, kV, c, cl, mag);
                pipe_beat_fast<1, LHY, MODE, PH>(B, a, fB, b, kH, kV, c, cl, mag);
                if (t - 7 >= y_out_lo) emit(t - 7, b);
                if (t + 3 < y_hi) load_row(r_in_re, r_in_im, t + 3, inB);
                load_e(t + 1, eB);
                load_e(t - 3, fB);
            } else {
                pair_general(t);
            }
        }
    } else {
        // FP64: the steady state in a loop of its own, whose simpler control
        // flow lets the compiler fit the pass in 215 VGPRs (2 waves per SIMD;
        // the single loop needed 256 plus spills to AGPRs, 1 wave per SIMD).
        // The iteration ends by waiting for its loads: the prefetched rows are
        // copied between registers at the back edge, which waits for them
        // anyway, and the early wait keeps the compiler from placing partial
        // waits inside the next beats (measured faster).
        if (interior) {
            for (; t < t_end && t < y_lo + 8; t += 2) pair_general(t);   // the pipelines fill
            for (; t + 1 < y_hi; t += 2) {
                pair_fast(t);
                __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
            }
        }
        for (; t < t_end; t += 2) pair_general(t);
    }
}

// FP32 with E as one phase plane and one scaled form on both axes: at most 128
// VGPRs (4 waves per SIMD; the compiler fits them without spills).  The host
// sizes the segments for the occupancy of the instance it launches.
#ifndef TSMI_STEP2_OCC
#define TSMI_STEP2_OCC 4
#endif
// One wave's share of a two-step pass: the FP32 stream where it applies (an
// interior segment of an interior strip), step2_wave otherwise.
template <typename T, int VEC, bool LHY, int MODE, bool PH, int SC = 0>
__device__ __forceinline__ void step2_item(const StepParams<T> &p, const int wave, const int lane,
                                           const int block_wave0 = -1) {
    if constexpr (TSMI_STREAM && sizeof(T) == 4 && VEC == 2 && PH && MODE >= 0) {
        constexpr int AP = 2 * kApron;
        constexpr int SW = kWave * VEC - 2 * AP;
        // interior strip and segment: the stream form
        auto stream_ok = [&](int w) {
            const int st = w % p.n_strips, sg = w / p.n_strips;
            const int x0 = st * SW - AP, lo = sg * p.seg_rows;
            return w < p.n_waves && x0 >= 0 && x0 + kWave * VEC <= p.W && lo - AP > 0 &&
                   min(p.H, lo + p.seg_rows) + AP < p.H;
        };
        const int strip = wave % p.n_strips;
        const int seg = wave / p.n_strips;
        const int X0 = strip * SW - AP;
        const int y_out_lo = seg * p.seg_rows;
        const int y_out_hi = min(p.H, y_out_lo + p.seg_rows);
        const int y_lo = y_out_lo - AP;
        const int y_hi = y_out_hi + AP;
        if (stream_ok(wave)) {
            // the workgroup's waves (neighbouring strips) keep step with a
            // barrier every kStreamSync beats when all of them run the stream
            // (the same segment height, so the same number of barriers)
            bool sync = false;
            if constexpr (kStreamSync > 0) {
                sync = block_wave0 >= 0;
#pragma unroll
                for (int i = 0; i < kStep2Block / kWave; ++i) sync = sync && stream_ok(block_wave0 + i);
            }
            const int xl = X0 + lane * VEC;
            const bool own = lane >= AP / VEC && lane < kWave - AP / VEC;
            const long seg0 = (long)y_lo * p.pitch;
            const unsigned seg_bytes = (unsigned)((long)(y_hi - y_lo) * p.pitch * (long)sizeof(T));
            const unsigned v_ld = (unsigned)xl * (unsigned)sizeof(T);
            const unsigned v_st = own ? v_ld : kOffNone;
            const T kH = MODE == 0 ? p.tH : p.mH, kV = MODE == 0 ? p.tV : p.mV;
            step2_stream_f32<LHY, MODE>(make_rsrc(p.in_re + seg0, seg_bytes), make_rsrc(p.in_im + seg0, seg_bytes),
                                        make_rsrc(p.ph + seg0, seg_bytes), make_rsrc(p.out_re + seg0, seg_bytes),
                                        make_rsrc(p.out_im + seg0, seg_bytes), v_ld, v_st,
                                        (unsigned)p.pitch * (unsigned)sizeof(T), y_hi - y_lo, AP,
                                        AP + (y_out_hi - y_out_lo), kH, kV, p.c_nl, p.c_lhy, p.mag, sync);
            return;
        }
    }
    step2_wave<T, VEC, LHY, MODE, PH, SC>(p, wave, lane);
}

// Waves per SIMD the FP32 phase-plane two-step instances are bounded to (128
// VGPRs: the stream's register slots without spills).
constexpr int step2_occupancy(int esz, bool ph, int mode) { return (esz == 4 && ph && mode >= 0) ? TSMI_STEP2_OCC : 1; }

// NT: output rows stored non-temporal (see kStoreNT).
template <typename T, int VEC, bool LHY, int MODE, bool PH = false, bool NT = false>
__global__ void __launch_bounds__(kStep2Block, step2_occupancy(sizeof(T), PH, MODE)) step2_kernel(StepParams<T> p) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave0 = xcd_block<sizeof(T) == 4>(blockIdx.x, gridDim.x) * (blockDim.x / kWave);
    const int wave = __builtin_amdgcn_readfirstlane(wave0 + (threadIdx.x / kWave));
    if (wave >= p.n_waves) return;
    step2_item<T, VEC, LHY, MODE, PH, NT ? kStoreNT : 0>(p, wave, lane, wave0);
}

}  // namespace tsmi
