/*
 * ts_oracle.c -- CPU restatement of the reference Trotter-Suzuki step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The reference (CPUBlock, src/cpukernel.cpp) evolves the tile in 128x128 cache
 * blocks that overlap by a 4-cell halo and keeps only each block's inner part.
 * One step's dependency cone is 4 cells (4 vertical and 4 horizontal nearest-
 * neighbour passes), and every block starts at a multiple of 120 (even), so the
 * blocked result equals one sweep of each pass over the whole tile with closed
 * tile edges.  This file states that whole-tile sweep directly; the arithmetic of
 * each pair update and of the potential step keeps the reference's expression
 * order so that, built without FMA contraction, it reproduces the reference
 * bit for bit (checked against the golden vectors).
 *
 * Reference defect NOT reproduced: for tile widths W > 128 with (W-128) % 120 == 0
 * (e.g. 248, 2048) CPUBlock's last column block starts 120 columns too far right
 * (src/cpukernel.cpp:111 versus the inner loop at :151) and the columns
 * [W-124, W-4) of the new buffer are never written.  The oracle evolves them.
 */
#include "ts_oracle.h"
#include <complex.h>
#include <math.h>
#include <stddef.h>

/* Pair update of two cells p, q by the real-time kinetic factor
 * [[a, i b], [i b, a]] (src/cpucartesian.cpp:6-11). */
static inline void pair_real(double a, double b, double *re, double *im, size_t p, size_t q) {
    const double pr = re[p], pi = im[p];
    re[p] = a * pr - b * im[q];
    im[p] = a * pi + b * re[q];
    re[q] = a * re[q] - b * pi;
    im[q] = a * im[q] + b * pr;
}

/* Pair update by the imaginary-time factor [[a, b], [b, a]]
 * (src/cpucartesian.cpp:27-32). */
static inline void pair_imag(double a, double b, double *re, double *im, size_t p, size_t q) {
    const double pr = re[p], pi = im[p];
    re[p] = a * pr + b * re[q];
    im[p] = a * pi + b * im[q];
    re[q] = a * re[q] + b * pr;
    im[q] = a * im[q] + b * pi;
}

void ts_oracle_coefficients(double dt, double mass, double dx, double dy,
                            double g, double g_lhy, int imag_time, double out[6]) {
    /* src/cpukernel.cpp:195-216 */
    const double bx = dt / (4. * mass * dx * dx);
    const double by = dt / (4. * mass * dy * dy);
    if (imag_time) {
        out[0] = cosh(bx); out[1] = sinh(bx);
        out[2] = cosh(by); out[3] = sinh(by);
    } else {
        out[0] = cos(bx); out[1] = sin(bx);
        out[2] = cos(by); out[3] = sin(by);
    }
    out[4] = g * dt;
    out[5] = g_lhy * dt;
}

void ts_oracle_exp_potential(long long n, const double *v, double dt, int imag_time,
                             double *e_re, double *e_im) {
    for (long long k = 0; k < n; ++k) {
        const double complex e = imag_time ? cexp(CMPLX(-dt * v[k], 0.)) : cexp(CMPLX(0., -dt * v[k]));
        e_re[k] = creal(e);
        e_im[k] = cimag(e);
    }
}

/* Tiles below this many cells run single-threaded: a parallel region costs more
 * than the pass (and much more when the host is shared). */
#define PAR_CELLS 65536

/* Vertical pass V(s): cells (x,y),(x,y+1) pair when x+y+s is even, y+1 < H
 * (src/cpucartesian.cpp:4-44). */
void ts_oracle_pass_vertical(const ts_oracle_desc *d, int s, double a, double b, double *re, double *im) {
    const size_t W = (size_t)d->W;
    const int H = d->H;
    #pragma omp parallel for schedule(static) if ((size_t)H * W >= PAR_CELLS)
    for (int y = 0; y < H - 1; ++y) {
        const size_t row = (size_t)y * W;
        for (size_t x = (size_t)((y + s) & 1); x < W; x += 2) {
            if (d->imag_time) pair_imag(a, b, re, im, row + x, row + W + x);
            else              pair_real(a, b, re, im, row + x, row + W + x);
        }
    }
}

/* Horizontal pass H(s): cells (x,y),(x+1,y) pair when x+y+s is even, x+1 < W
 * (src/cpucartesian.cpp:46-70). */
void ts_oracle_pass_horizontal(const ts_oracle_desc *d, int s, double a, double b, double *re, double *im) {
    const size_t W = (size_t)d->W;
    const int H = d->H;
    #pragma omp parallel for schedule(static) if ((size_t)H * W >= PAR_CELLS)
    for (int y = 0; y < H; ++y) {
        const size_t row = (size_t)y * W;
        for (size_t x = (size_t)((y + s) & 1); x + 1 < W; x += 2) {
            if (d->imag_time) pair_imag(a, b, re, im, row + x, row + x + 1);
            else              pair_real(a, b, re, im, row + x, row + x + 1);
        }
    }
}

/* Potential and nonlinear phase P (src/cpucartesian.cpp:93-109 real time,
 * :127-137 imaginary time).  |psi|^2 is taken before the multiplication by E. */
void ts_oracle_pass_potential(const ts_oracle_desc *d, double *re, double *im,
                              const double *e_re, const double *e_im) {
    const size_t n = (size_t)d->W * (size_t)d->H;
    const double c = d->c_nl, cl = d->c_lhy;
    #pragma omp parallel for schedule(static) if (n >= PAR_CELLS)
    for (long long k = 0; k < (long long)n; ++k) {
        const double r = re[k], i = im[k];
        const double n2 = r * r + i * i;
        const double n3 = n2 * sqrt(n2);
        if (d->imag_time) {
            const double f = exp(-1. * (c * n2 + cl * n3));
            re[k] = f * e_re[k] * r;
            im[k] = f * e_re[k] * i;
        } else {
            const double ang = c * n2 + cl * n3;
            const double cs = cos(ang), sn = sin(ang);
            const double r1 = e_re[k] * r - e_im[k] * i;
            const double i1 = e_re[k] * i + e_im[k] * r;
            re[k] = cs * r1 + sn * i1;
            im[k] = cs * i1 - sn * r1;
        }
    }
}

void ts_oracle_full_step(const ts_oracle_desc *d, double *re, double *im,
                         const double *e_re, const double *e_im) {
    /* src/cpukernel.cpp:22-56; vertical passes only when the block has >1 row. */
    const int vert = d->H > 1;
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 1, d->aH, d->bH, re, im);
    ts_oracle_pass_potential(d, re, im, e_re, e_im);
    ts_oracle_pass_horizontal(d, 1, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, d->bV, re, im);
}

/* Element-wise forward copy of a width x height sub-matrix inside one plane. */
static void copy2d(double *p, size_t dst, size_t src, size_t stride, size_t width, size_t height) {
    for (size_t i = 0; i < height; ++i)
        for (size_t j = 0; j < width; ++j)
            p[dst + i * stride + j] = p[src + i * stride + j];
}

void ts_oracle_halo_exchange(const ts_oracle_desc *d, double *re, double *im) {
    const size_t W = (size_t)d->W, H = (size_t)d->H;
    const size_t hx = (size_t)d->halo_x, hy = (size_t)d->halo_y;
    if (d->periodic_x) {
        /* src/cpukernel.cpp:708-714: rows from the first inner row, count
         * H - 2*halo_y (halo_y is the lattice halo even on a closed y axis). */
        const size_t iy0 = d->periodic_y ? hy : 0;
        const size_t rows = H > 2 * hy ? H - 2 * hy : 0;
        const size_t off = iy0 * W;
        copy2d(re, off, off + W - 2 * hx, W, hx, rows);
        copy2d(im, off, off + W - 2 * hx, W, hx, rows);
        copy2d(re, off + W - hx, off + hx, W, hx, rows);
        copy2d(im, off + W - hx, off + hx, W, hx, rows);
    }
    if (d->periodic_y) {
        /* src/cpukernel.cpp:739-745: full-width rows. */
        const size_t off = (H - hy) * W;
        copy2d(re, 0, off - hy * W, W, W, hy);
        copy2d(im, 0, off - hy * W, W, W, hy);
        copy2d(re, off, hy * W, W, W, hy);
        copy2d(im, off, hy * W, W, W, hy);
    }
}

static double inner_sum(const ts_oracle_desc *d, const double *re, const double *im) {
    const size_t W = (size_t)d->W;
    const int x0 = d->periodic_x ? d->halo_x : 0, x1 = d->W - x0;
    const int y0 = d->periodic_y ? d->halo_y : 0, y1 = d->H - y0;
    double s = 0.;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            const size_t k = (size_t)y * W + (size_t)x;
            s += re[k] * re[k] + im[k] * im[k];
        }
    return s;
}

double ts_oracle_squared_norm(const ts_oracle_desc *d, const double *re, const double *im) {
    /* src/cpukernel.cpp:526-549 over the inner region, scaled as (s*dx)*dy. */
    return inner_sum(d, re, im) * d->dx * d->dy;
}

double ts_oracle_state_norm(const ts_oracle_desc *d, const double *re, const double *im) {
    return inner_sum(d, re, im) * (d->dx * d->dy);
}

void ts_oracle_evolve(const ts_oracle_desc *d, int n_steps, double *re, double *im,
                      const double *e_re, const double *e_im) {
    const size_t n = (size_t)d->W * (size_t)d->H;
    for (int i = 0; i < n_steps; ++i) {
        ts_oracle_full_step(d, re, im, e_re, e_im);
        if (i != n_steps - 1)
            ts_oracle_halo_exchange(d, re, im);          /* src/solver.cpp:203-209 */
        if (d->imag_time && d->norm_target != 0) {       /* src/cpukernel.cpp:552-563 */
            const double f = sqrt(ts_oracle_squared_norm(d, re, im) / d->norm_target);
            for (size_t k = 0; k < n; ++k) {
                re[k] /= f;
                im[k] /= f;
            }
        }
    }
}

/* Inner-region bounds and the 2-cell stencil margin next to closed edges
 * (src/solver.cpp:265-269,356-359; src/model.cpp:386-390,424-427). */
typedef struct { int x0, x1, y0, y1, kx0, kx1, ky0, ky1; } region;

static region inner_region(const ts_oracle_desc *d) {
    region r;
    r.x0 = d->periodic_x ? d->halo_x : 0; r.x1 = d->W - r.x0;
    r.y0 = d->periodic_y ? d->halo_y : 0; r.y1 = d->H - r.y0;
    r.kx0 = r.x0 + (d->periodic_x ? 0 : 2); r.kx1 = r.x1 - (d->periodic_x ? 0 : 2);
    r.ky0 = r.y0 + (d->periodic_y ? 0 : 2); r.ky1 = r.y1 - (d->periodic_y ? 0 : 2);
    return r;
}

void ts_oracle_energy(const ts_oracle_desc *d, const double *re, const double *im,
                      const double *v, double mass, double g, double g_lhy,
                      double Lx, int Nx, double out[6]) {
    const size_t W = (size_t)d->W;
    const region r = inner_region(d);
    const double cost = -1. / (2. * mass);
    const double c1 = -1. / 12., c2 = 4. / 3., c3 = -2.5;
    const double idx2 = 1. / (d->dx * d->dx), idy2 = 1. / (d->dy * d->dy);
    double s_n = 0, s_nk = 0, s_kin = 0, s_pot = 0, s_int = 0, s_lhy = 0;
    for (int y = r.y0; y < r.y1; ++y) {
        for (int x = r.x0; x < r.x1; ++x) {
            const size_t k = (size_t)y * W + (size_t)x;
            const double pr = re[k], pi = im[k];
            const double n2 = pr * pr + pi * pi;
            s_n += n2;
            s_pot += n2 * v[k];
            s_int += n2 * n2 * 0.5 * g;
            s_lhy += pow(n2, 2.5) * 0.4 * g_lhy;
            const int in_margin = ((y >= r.ky0 && y < r.ky1) || d->H == 1) && (x >= r.kx0 && x < r.kx1);
            if (!in_margin) continue;
            s_nk += n2;
            /* 4th-order Laplacian, src/solver.cpp:379-383 */
            const double lxr = c1 * re[k + 2] + c2 * re[k + 1] + c2 * re[k - 1] + c1 * re[k - 2] + c3 * pr;
            const double lxi = c1 * im[k + 2] + c2 * im[k + 1] + c2 * im[k - 1] + c1 * im[k - 2] + c3 * pi;
            s_kin += cost * (pr * lxr + pi * lxi) * idx2;
            if (d->H > 1) {
                const double lyr = c1 * re[k + 2 * W] + c2 * re[k + W] + c2 * re[k - W] + c1 * re[k - 2 * W] + c3 * pr;
                const double lyi = c1 * im[k + 2 * W] + c2 * im[k + W] + c2 * im[k - W] + c1 * im[k - 2 * W] + c3 * pi;
                s_kin += cost * (pr * lyr + pi * lyi) * idy2;
            }
        }
    }
    const double kin = s_kin / s_nk, pot = s_pot / s_n, itr = s_int / s_n, lhy = s_lhy / s_n;
    out[0] = kin + pot + itr + lhy;        /* src/solver.cpp:538 (rotation term is 0) */
    out[1] = kin;
    out[2] = pot;
    out[3] = itr;
    out[4] = lhy;
    out[5] = s_n * d->dy * Lx / (double)Nx; /* src/solver.cpp:561 */
}

void ts_oracle_expected_values(const ts_oracle_desc *d, const double *re, const double *im,
                               const double *xc, const double *yc, double out[10]) {
    const size_t W = (size_t)d->W;
    const region r = inner_region(d);
    const double c1 = -1. / 12., c2 = 4. / 3., c3 = -2.5;
    const double d1 = 1. / 6., d2 = -1., d3 = 0.5, d4 = 1. / 3.;
    double sn = 0, sx = 0, sxx = 0, sy = 0, syy = 0, spx = 0, spy = 0, spxpx = 0, spypy = 0, slz = 0;
    for (int y = r.y0; y < r.y1; ++y) {
        for (int x = r.x0; x < r.x1; ++x) {
            const size_t k = (size_t)y * W + (size_t)x;
            const double pr = re[k], pi = im[k];
            const double n2 = pr * pr + pi * pi;
            const double X = xc[x], Y = yc[y];
            sn += n2; sx += n2 * X; sy += n2 * Y; sxx += n2 * X * X; syy += n2 * Y * Y;
            if (!(y >= r.ky0 && y < r.ky1 && x >= r.kx0 && x < r.kx1)) continue;
            /* one-sided first derivative, src/model.cpp:447-448 (up = row y-1) */
            const double dxr = d4 * re[k + 1] + d3 * pr + d2 * re[k - 1] + d1 * re[k - 2];
            const double dxi = d4 * im[k + 1] + d3 * pi + d2 * im[k - 1] + d1 * im[k - 2];
            const double dyr = d4 * re[k - W] + d3 * pr + d2 * re[k + W] + d1 * re[k + 2 * W];
            const double dyi = d4 * im[k - W] + d3 * pi + d2 * im[k + W] + d1 * im[k + 2 * W];
            /* imag(conj(psi) * z) = pr*z_im - pi*z_re */
            spx += pr * dxi - pi * dxr;
            spy += pr * dyi - pi * dyr;
            const double lxr = c1 * re[k + 2] + c2 * re[k + 1] + c2 * re[k - 1] + c1 * re[k - 2] + c3 * pr;
            const double lxi = c1 * im[k + 2] + c2 * im[k + 1] + c2 * im[k - 1] + c1 * im[k - 2] + c3 * pi;
            const double lyr = c1 * re[k + 2 * W] + c2 * re[k + W] + c2 * re[k - W] + c1 * re[k - 2 * W] + c3 * pr;
            const double lyi = c1 * im[k + 2 * W] + c2 * im[k + W] + c2 * im[k - W] + c1 * im[k - 2 * W] + c3 * pi;
            spxpx += pr * lxr + pi * lxi;
            spypy += pr * lyr + pi * lyi;
            const double zr = Y / d->dx * dxr + X / d->dy * dyr;
            const double zi = Y / d->dx * dxi + X / d->dy * dyi;
            slz += pr * zi - pi * zr;
        }
    }
    const double ppx = -1. / d->dx, ppy = 1. / d->dy;   /* src/model.cpp:398-399 */
    out[0] = sn * d->dx * d->dy;
    out[1] = sx / sn;
    out[2] = sxx / sn;
    out[3] = sy / sn;
    out[4] = syy / sn;
    out[5] = -spx * ppx / sn;
    out[6] = -spxpx * ppx * ppx / sn;
    out[7] = -spy * ppy / sn;
    out[8] = -spypy * ppy * ppy / sn;
    out[9] = slz / sn;
}

/* ---------------------------------------------------------------------------
 * Discrete adjoint of the real-time evolution with respect to the potential.
 *
 * Not a reference function: the reference has no gradient.  It is the exact
 * derivative of ts_oracle_evolve in real time (any boundaries, any g and
 * g_LHY) with respect to every tile entry of V and of psi_0, so it is pinned by
 * central finite differences of that forward (and therefore of the reference
 * solver, which the forward matches bit for bit).
 *
 * Step i of the forward is psi_{i+1} = X_i A N(A' psi_i): A' = the first four
 * passes V0 H0 V1 H1, N the potential stage, A the last four, X_i the halo
 * exchange (after every step but the last).  With the adjoint state
 * G = dL/dRe psi + i dL/dIm psi, one reverse step is
 *     G <- X_i^T G,  mu = A^H G,  G <- A'^H N^T mu,
 * where a real-time pass [[a, ib], [ib, a]] is symmetric, so its adjoint is the
 * same pass with b -> -b and the passes of A^H run in the order of A'.  At a
 * cell with chi = A' psi_i, s = |chi|^2, phi = c s + c_LHY s^(3/2) and
 * y = N(chi) = E chi e^(-i phi) (the forward's own arithmetic):
 *     dL/dV += dt Im(conj(mu) y),
 *     N^T mu = conj(E e^(-i phi)) mu + 2 phi'(s) Im(conj(mu) y) chi,
 * phi'(s) = c + 1.5 c_LHY s^(1/2); with c = c_LHY = 0 this is conj(E) mu.  The
 * exchange is a sequence of copies dst <- src; the adjoint of each adds dst
 * into src and clears dst, in reverse order.  Forward states are kept every
 * `every` steps and each segment is recomputed from its checkpoint before its
 * reverse sweep.
 * ------------------------------------------------------------------------- */
#include <stdlib.h>
#include <string.h>

static void first_half(const ts_oracle_desc *d, double sgn, double *re, double *im) {
    const int vert = d->H > 1;
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, sgn * d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, sgn * d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, sgn * d->bV, re, im);
    ts_oracle_pass_horizontal(d, 1, d->aH, sgn * d->bH, re, im);
}

static void second_half(const ts_oracle_desc *d, double sgn, double *re, double *im) {
    const int vert = d->H > 1;
    ts_oracle_pass_horizontal(d, 1, d->aH, sgn * d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, sgn * d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, sgn * d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, sgn * d->bV, re, im);
}

/* Adjoint of copy2d: every element's copy undone in reverse order. */
static void copy2d_adjoint(double *p, size_t dst, size_t src, size_t stride, size_t width, size_t height) {
    for (size_t i = height; i-- > 0;)
        for (size_t j = width; j-- > 0;) {
            p[src + i * stride + j] += p[dst + i * stride + j];
            p[dst + i * stride + j] = 0.;
        }
}

void ts_oracle_halo_exchange_adjoint(const ts_oracle_desc *d, double *re, double *im) {
    /* the copies of ts_oracle_halo_exchange, last first */
    const size_t W = (size_t)d->W, H = (size_t)d->H;
    const size_t hx = (size_t)d->halo_x, hy = (size_t)d->halo_y;
    if (d->periodic_y) {
        const size_t off = (H - hy) * W;
        copy2d_adjoint(im, off, hy * W, W, W, hy);
        copy2d_adjoint(re, off, hy * W, W, W, hy);
        copy2d_adjoint(im, 0, off - hy * W, W, W, hy);
        copy2d_adjoint(re, 0, off - hy * W, W, W, hy);
    }
    if (d->periodic_x) {
        const size_t iy0 = d->periodic_y ? hy : 0;
        const size_t rows = H > 2 * hy ? H - 2 * hy : 0;
        const size_t off = iy0 * W;
        copy2d_adjoint(im, off + W - hx, off + hx, W, hx, rows);
        copy2d_adjoint(re, off + W - hx, off + hx, W, hx, rows);
        copy2d_adjoint(im, off, off + W - 2 * hx, W, hx, rows);
        copy2d_adjoint(re, off, off + W - 2 * hx, W, hx, rows);
    }
}

/* N^T at every cell (see above); chi is the forward's pre-potential state. */
static void potential_adjoint(const ts_oracle_desc *d, const double *chr, const double *chi_i,
                              const double *e_re, const double *e_im, double dt,
                              double *gr, double *gi, double *grad_v) {
    const size_t n = (size_t)d->W * (size_t)d->H;
    const double c = d->c_nl, cl = d->c_lhy;
    for (size_t k = 0; k < n; ++k) {
        const double cr = chr[k], ci = chi_i[k];
        const double er = e_re[k], ei = e_im[k];
        const double n2 = cr * cr + ci * ci;
        const double ang = c * n2 + cl * (n2 * sqrt(n2));
        const double cs = cos(ang), sn = sin(ang);
        const double r1 = er * cr - ei * ci, i1 = er * ci + ei * cr;   /* E chi */
        const double yr = cs * r1 + sn * i1, yi = cs * i1 - sn * r1;    /* y */
        const double mr = gr[k], mi = gi[k];
        const double z = mr * yi - mi * yr;                             /* Im(conj(mu) y) */
        grad_v[k] += dt * z;
        const double ar = er * mr + ei * mi, ai = er * mi - ei * mr;    /* conj(E) mu */
        const double f = 2. * (c + 1.5 * cl * sqrt(n2)) * z;
        gr[k] = (cs * ar - sn * ai) + f * cr;
        gi[k] = (sn * ar + cs * ai) + f * ci;
    }
}

int ts_oracle_adjoint(const ts_oracle_desc *d, int n_steps, int every,
                      const double *psi_re, const double *psi_im,
                      const double *e_re, const double *e_im, double dt,
                      const double *g_re, const double *g_im,
                      double *grad_v, double *g0_re, double *g0_im) {
    if (d->imag_time || n_steps < 0 || every < 1)
        return -1;
    const size_t n = (size_t)d->W * (size_t)d->H;
    const int n_ck = n_steps > 0 ? (n_steps - 1) / every + 1 : 0;
    const int seg_len = n_steps < every ? n_steps : every;
    double *ck = malloc(sizeof(double) * 2 * n * (size_t)(n_ck > 0 ? n_ck : 1));
    double *chi = malloc(sizeof(double) * 2 * n * (size_t)(seg_len > 0 ? seg_len : 1));
    double *st = malloc(sizeof(double) * 2 * n);
    double *gg = malloc(sizeof(double) * 2 * n);
    if (!ck || !chi || !st || !gg) {
        free(ck); free(chi); free(st); free(gg);
        return -2;
    }
    /* forward sweep, keeping psi_{k*every} */
    memcpy(st, psi_re, sizeof(double) * n);
    memcpy(st + n, psi_im, sizeof(double) * n);
    for (int i = 0; i < n_steps; ++i) {
        if (i % every == 0) memcpy(ck + 2 * n * (size_t)(i / every), st, sizeof(double) * 2 * n);
        ts_oracle_full_step(d, st, st + n, e_re, e_im);
        if (i != n_steps - 1) ts_oracle_halo_exchange(d, st, st + n);
    }
    memcpy(gg, g_re, sizeof(double) * n);
    memcpy(gg + n, g_im, sizeof(double) * n);
    for (size_t k = 0; k < n; ++k) grad_v[k] = 0.;
    double *gr = gg, *gi = gg + n;
    for (int seg = n_ck - 1; seg >= 0; --seg) {
        const int s0 = seg * every;
        const int s1 = s0 + every < n_steps ? s0 + every : n_steps;
        /* recompute chi_i = A' psi_i over the segment */
        memcpy(st, ck + 2 * n * (size_t)seg, sizeof(double) * 2 * n);
        for (int i = s0; i < s1; ++i) {
            first_half(d, 1., st, st + n);
            memcpy(chi + 2 * n * (size_t)(i - s0), st, sizeof(double) * 2 * n);
            ts_oracle_pass_potential(d, st, st + n, e_re, e_im);
            second_half(d, 1., st, st + n);
            if (i != n_steps - 1) ts_oracle_halo_exchange(d, st, st + n);
        }
        for (int i = s1 - 1; i >= s0; --i) {
            const double *cr = chi + 2 * n * (size_t)(i - s0), *ci = cr + n;
            if (i != n_steps - 1) ts_oracle_halo_exchange_adjoint(d, gr, gi);
            first_half(d, -1., gr, gi);                      /* mu = A^H G */
            potential_adjoint(d, cr, ci, e_re, e_im, dt, gr, gi, grad_v);
            second_half(d, -1., gr, gi);                     /* A'^H */
        }
    }
    if (g0_re) memcpy(g0_re, gr, sizeof(double) * n);
    if (g0_im) memcpy(g0_im, gi, sizeof(double) * n);
    free(ck); free(chi); free(st); free(gg);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Two-component systems
 * ------------------------------------------------------------------------- */

/* Potential pass of one component of two (src/cpucartesian.cpp:75-92 real
 * time, :115-126 imaginary time): the phase sees the component's own density
 * and the other component's density pb, taken from the start of the step. */
static void pass_potential2(const ts_oracle_desc *d, double *re, double *im, const double *e_re,
                            const double *e_im, const double *pb_re, const double *pb_im, double c_ab) {
    const size_t n = (size_t)d->W * (size_t)d->H;
    const double c = d->c_nl;
    #pragma omp parallel for schedule(static) if (n >= PAR_CELLS)
    for (long long k = 0; k < (long long)n; ++k) {
        const double r = re[k], i = im[k];
        const double n2 = r * r + i * i;
        const double n2b = pb_re[k] * pb_re[k] + pb_im[k] * pb_im[k];
        if (d->imag_time) {
            const double f = exp(-1. * (c * n2 + c_ab * n2b));
            re[k] = f * e_re[k] * r;
            im[k] = f * e_re[k] * i;
        } else {
            const double ang = c * n2 + c_ab * n2b;
            const double cs = cos(ang), sn = sin(ang);
            const double r1 = e_re[k] * r - e_im[k] * i;
            const double i1 = e_re[k] * i + e_im[k] * r;
            re[k] = cs * r1 + sn * i1;
            im[k] = cs * i1 - sn * r1;
        }
    }
}

/* full_step / full_step_imaginary with two wave functions (src/cpukernel.cpp:22-92). */
static void full_step2(const ts_oracle_desc *d, double *re, double *im, const double *e_re, const double *e_im,
                       const double *pb_re, const double *pb_im, double c_ab) {
    const int vert = d->H > 1;
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 1, d->aH, d->bH, re, im);
    pass_potential2(d, re, im, e_re, e_im, pb_re, pb_im, c_ab);
    ts_oracle_pass_horizontal(d, 1, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 1, d->aV, d->bV, re, im);
    ts_oracle_pass_horizontal(d, 0, d->aH, d->bH, re, im);
    if (vert) ts_oracle_pass_vertical(d, 0, d->aV, d->bV, re, im);
}

void ts_oracle_rabi(const ts_oracle_desc2 *d, double var, double dt,
                    double *a_re, double *a_im, double *b_re, double *b_im) {
    /* src/cpukernel.cpp:581-608; the constants are half the Hamiltonian's omega (:330-331) */
    const double cr = 0.5 * d->omega_r, ci = 0.5 * d->omega_i;
    const double norm_omega = sqrt(cr * cr + ci * ci);
    const int imag = d->a.imag_time;
    double cc, cs_r, cs_i;
    if (imag) {
        cc = cosh(-dt * var * norm_omega);
        cs_r = norm_omega == 0 ? 0 : cr / norm_omega * sinh(-dt * var * norm_omega);
        cs_i = norm_omega == 0 ? 0 : ci / norm_omega * sinh(-dt * var * norm_omega);
    } else {
        cc = cos(-dt * var * norm_omega);
        cs_r = norm_omega == 0 ? 0 : cr / norm_omega * sin(-dt * var * norm_omega);
        cs_i = norm_omega == 0 ? 0 : ci / norm_omega * sin(-dt * var * norm_omega);
    }
    const size_t n = (size_t)d->a.W * (size_t)d->a.H;
    for (size_t k = 0; k < n; ++k) {
        const double r = a_re[k], i = a_im[k];
        if (imag) {   /* src/cpucartesian.cpp:278-290 */
            a_re[k] = cc * r + cs_r * b_re[k] - cs_i * b_im[k];
            a_im[k] = cc * i + cs_i * b_re[k] + cs_r * b_im[k];
            b_re[k] = cc * b_re[k] + cs_r * r + cs_i * i;
            b_im[k] = cc * b_im[k] - cs_i * r + cs_r * i;
        } else {      /* src/cpucartesian.cpp:264-276 */
            a_re[k] = cc * r - cs_i * b_re[k] - cs_r * b_im[k];
            a_im[k] = cc * i + cs_r * b_re[k] - cs_i * b_im[k];
            b_re[k] = cc * b_re[k] + cs_i * r - cs_r * i;
            b_im[k] = cc * b_im[k] + cs_r * r + cs_i * i;
        }
    }
}

static void scale_tile(size_t n, double f, double *re, double *im) {
    for (size_t k = 0; k < n; ++k) {
        re[k] /= f;
        im[k] /= f;
    }
}

void ts_oracle_evolve2(const ts_oracle_desc2 *d, int n_steps, int initial_half, double dt,
                       double *a_re, double *a_im, double *b_re, double *b_im,
                       const double *ea_re, const double *ea_im, const double *eb_re, const double *eb_im,
                       double norm[2], double tot_norm) {
    const size_t n = (size_t)d->a.W * (size_t)d->a.H;
    const int imag = d->a.imag_time;
    double *old_re = (double *)malloc(n * sizeof(double)), *old_im = (double *)malloc(n * sizeof(double));
    if (!old_re || !old_im) {
        free(old_re);
        free(old_im);
        return;
    }
    if (initial_half) ts_oracle_rabi(d, 0.5, dt, a_re, a_im, b_re, b_im);
    for (int i = 0; i < n_steps; ++i) {
        /* component b's step reads the start-of-step a (p[0][sense] in
         * src/cpukernel.cpp:421-424 after wait_for_completion flipped sense back) */
        memcpy(old_re, a_re, n * sizeof(double));
        memcpy(old_im, a_im, n * sizeof(double));
        full_step2(&d->a, a_re, a_im, ea_re, ea_im, b_re, b_im, d->c_ab);
        if (i != n_steps - 1) ts_oracle_halo_exchange(&d->a, a_re, a_im);
        if (imag && norm[0] != 0) scale_tile(n, sqrt(ts_oracle_squared_norm(&d->a, a_re, a_im) / norm[0]), a_re, a_im);
        full_step2(&d->b, b_re, b_im, eb_re, eb_im, old_re, old_im, d->c_ab);
        if (i != n_steps - 1) ts_oracle_halo_exchange(&d->b, b_re, b_im);
        if (imag && norm[1] != 0) scale_tile(n, sqrt(ts_oracle_squared_norm(&d->b, b_re, b_im) / norm[1]), b_re, b_im);
        ts_oracle_rabi(d, i == n_steps - 1 ? 0.5 : 1.0, dt, a_re, a_im, b_re, b_im);
        if (imag && (d->omega_r != 0 || d->omega_i != 0)) {
            /* src/cpukernel.cpp:610-667 */
            const double sa = inner_sum(&d->a, a_re, a_im), sb = inner_sum(&d->b, b_re, b_im);
            const double f = sqrt((sa + sb) * d->a.dx * d->a.dy / tot_norm);
            scale_tile(n, f, a_re, a_im);
            norm[0] = sa / (sa + sb) * tot_norm;
            scale_tile(n, f, b_re, b_im);
            norm[1] = sb / (sa + sb) * tot_norm;
        }
    }
    free(old_re);
    free(old_im);
}

void ts_oracle_energy2(const ts_oracle_desc2 *d, const double *a_re, const double *a_im,
                       const double *b_re, const double *b_im, const double *va, const double *vb,
                       double mass_a, double mass_b, double g_a, double g_b, double g_ab,
                       double omega_r, double omega_i, double Lx, int Nx, double out[12]) {
    double ea[6], eb[6];
    ts_oracle_energy(&d->a, a_re, a_im, va, mass_a, g_a, 0.0, Lx, Nx, ea);
    ts_oracle_energy(&d->b, b_re, b_im, vb, mass_b, g_b, 0.0, Lx, Nx, eb);
    /* the cross sums and the plain norms (src/solver.cpp:340-354, :546-550) */
    const size_t W = (size_t)d->a.W;
    const region r = inner_region(&d->a);
    const double complex om = CMPLX(omega_r, omega_i);
    double s_a = 0, s_b = 0, s_ab = 0, s_rabi = 0;
    for (int y = r.y0; y < r.y1; ++y)
        for (int x = r.x0; x < r.x1; ++x) {
            const size_t k = (size_t)y * W + (size_t)x;
            const double complex pa = CMPLX(a_re[k], a_im[k]), pb = CMPLX(b_re[k], b_im[k]);
            s_a += creal(conj(pa) * pa);
            s_b += creal(conj(pb) * pb);
            s_ab += creal(conj(pa) * pa * conj(pa) * pa * conj(pb) * pb * conj(pb) * pb * g_ab);
            s_rabi += creal(conj(pa) * pa * conj(pb) * pb * (conj(pa) * pb * om + conj(pb * om) * pa));
        }
    const double inter = s_ab / (s_a * s_b), rabi = 0.5 * s_rabi / (s_a * s_b);
    out[1] = ea[1]; out[2] = eb[1];
    out[3] = ea[2]; out[4] = eb[2];
    out[5] = ea[3]; out[6] = eb[3];
    out[7] = inter; out[8] = rabi;
    out[0] = ea[1] + ea[2] + ea[3] + eb[1] + eb[2] + eb[3] + inter + rabi;   /* src/solver.cpp:552-554 */
    out[9] = ea[5]; out[10] = eb[5];
    out[11] = ea[5] + eb[5];
}
