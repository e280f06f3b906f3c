// frame.hpp -- host-side scaled-frame factors of the pair updates (see the
// note at fpair in step_kernel.hpp).
#pragma once
#include <cmath>
#include <complex>

#include "tsmi_hip.h"

namespace tsmi {

// Scaled-frame factors of one axis: form A (m, t) = (1, b/a) when |a| >= |b| or
// in imaginary time, form B (a/b, 1) otherwise, both rounded to the kernel
// precision.  g = lambda / lambda~ compares the exact smooth-mode eigenvalue of
// a pass (a + ib, or a + b in imaginary time) with the computed one (m + it, or
// m + t); its modulus is exact for every mode.  A cell without a partner in a
// pass is multiplied by un = 1/Re(g).
struct Axis { double m, t, un; std::complex<double> g; int form; };

inline Axis frame_axis(double a, double b, bool imag, int esz) {
    Axis x;
    x.form = (imag || std::fabs(a) >= std::fabs(b)) ? 0 : 1;
    double m = x.form == 0 ? 1.0 : a / b, t = x.form == 0 ? b / a : 1.0;
    if (esz == 4) { m = (float)m; t = (float)t; }
    x.m = m; x.t = t;
    const std::complex<double> lam = imag ? std::complex<double>(a + b, 0.0) : std::complex<double>(a, b);
    const std::complex<double> lt = imag ? std::complex<double>(m + t, 0.0) : std::complex<double>(m, t);
    x.g = lam / lt;
    double un = 1.0 / x.g.real();
    if (esz == 4) un = (float)un;
    x.un = un;
    return x;
}

// Everything a step launch needs from the descriptor, for a tile of H rows in
// the precision esz (4 or 8 bytes).
struct StepCoefs {
    double aH = 1, bH = 0, aV = 1, bV = 0;     // FP64, as the reference (src/cpukernel.cpp:195-216)
    double c_nl = 0, c_lhy = 0;
    double mH = 1, tH = 0, mV = 1, tV = 0, unH = 1, unV = 1;   // kernel frame
    double corr_re = 1, corr_im = 0;           // folded into E
    double c_nl_s = 0, c_lhy_s = 0;            // couplings in the mid-step frame
    double corr_abs = 1;                       // |corr|
    int mode = 0;                              // 0 / 1: both axes in form A / B; -1: mixed
    // frame of the step at -dt (the reverse sweep): b -> -b, same forms
    double amH = 1, atH = 0, amV = 1, atV = 0, aunH = 1, aunV = 1;
};

inline StepCoefs step_coefs(const tsmi_desc &d, int H, int esz) {
    StepCoefs k;
    const double bx = d.delta_t / (4. * d.mass * d.delta_x * d.delta_x);
    const double by = d.delta_t / (4. * d.mass * d.delta_y * d.delta_y);
    if (d.imag_time) {
        k.aH = std::cosh(bx); k.bH = std::sinh(bx);
        k.aV = std::cosh(by); k.bV = std::sinh(by);
    } else {
        k.aH = std::cos(bx); k.bH = std::sin(bx);
        k.aV = std::cos(by); k.bV = std::sin(by);
    }
    k.c_nl = d.coupling * d.delta_t;
    k.c_lhy = d.coupling_lhy * d.delta_t;
    // A step runs 4 passes per axis, 2 of them before the potential step, so E
    // takes corr = gH^4 gV^4 and the couplings see |psi|^2 / |gH gV|^4.
    const bool imag = d.imag_time != 0;
    const Axis one{1.0, 0.0, 1.0, std::complex<double>(1.0, 0.0), 0};
    const Axis ah = frame_axis(k.aH, k.bH, imag, esz);
    Axis av = one;
    if (H > 1) av = frame_axis(k.aV, k.bV, imag, esz);
    else av.form = ah.form;
    k.mH = ah.m; k.tH = ah.t; k.unH = ah.un;
    k.mV = av.m; k.tV = av.t; k.unV = av.un;
    k.mode = ah.form == av.form ? ah.form : -1;
    std::complex<double> corr = ah.g * ah.g * ah.g * ah.g;
    if (H > 1) corr *= av.g * av.g * av.g * av.g;
    k.corr_re = corr.real();
    k.corr_im = imag ? 0.0 : corr.imag();
    const double sn = std::pow(std::abs(ah.g) * std::abs(av.g), 4);
    k.corr_abs = sn;
    k.c_nl_s = k.c_nl * sn;
    k.c_lhy_s = k.c_lhy * std::pow(sn, 1.5);
    if (!imag) {
        const Axis bh = frame_axis(k.aH, -k.bH, false, esz);
        Axis bv = one;
        if (H > 1) bv = frame_axis(k.aV, -k.bV, false, esz);
        k.amH = bh.m; k.atH = bh.t; k.aunH = bh.un;
        k.amV = bv.m; k.atV = bv.t; k.aunV = bv.un;
    }
    return k;
}

}  // namespace tsmi
