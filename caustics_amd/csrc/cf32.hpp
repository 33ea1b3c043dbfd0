// Reduced-precision Ehrlich-Aberth form ("CF32"): FP32 iterations, one compensated
// FP32 sweep, FP64 polish.
//
//   stage A  plain Ehrlich-Aberth in FP32 on the coefficients scaled by a power of two
//            (max |p_k| -> [1,2)), forward polynomial, stopping test |h| <= 2^-24 b.
//   stage B  one sweep of compensated FP32 Horner (TwoProd by FMA, TwoSum; Ogita-Rump-
//            Oishi / Graillat-Langlois-Louvet compensated Horner) on double-float
//            coefficients (hi + lo FP32 words) evaluated at the FP32 iterate x; the Aberth
//            correction is formed in FP32 and applied in double-float, x - corr, so the
//            updated root carries ~2x FP32 precision.
//   stage C  the FP64 stopping test of the plain solver (ehrlich_aberth.h:100-143) on the
//            FP64 coefficients, iterating in FP64 until it passes (normally it passes on
//            the first evaluation).  The returned roots therefore satisfy the same FP64
//            criterion as the FP64 form.
// A non-finite FP32 iterate, or a non-finite result (two estimates that FP32 could not
// separate), sends the polynomial through the FP64 solver from scratch.
#pragma once
#include "solver.hpp"

namespace caus {

constexpr float EA_EPS32 = 5.9604644775390625e-08f;  // 2^-24

__device__ __forceinline__ void two_sum_f(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    float t = s - a;
    e = (a - (s - t)) + (b - t);
}
__device__ __forceinline__ void two_prod_f(float a, float b, float& p, float& e) {
#pragma clang fp contract(off)
    p = a * b;
    e = fmaf(a, b, -p);
}

template <int DEG>
__device__ __forceinline__ cf aberth_sum_f(const cf (&z)[DEG], int j, cf x) {
    cf S = cf{0.f, 0.f};
#pragma unroll
    for (int i = 0; i < DEG; ++i)
        if (i != j) S = S + crcp(x - z[i]);
    return S;
}

// GUARD: the public root solver's guarded FP64 stage (fast_step)
template <int DEG, bool GUARD = false>
__device__ __forceinline__ bool solve_cf32(const cd (&p)[DEG + 1], cd (&z)[DEG], int itmax,
                                           bool custom_init, SolveStats& st) {
    st.sweeps = 0;
    st.it64 = 0;
    st.itcomp = 0;
    st.it32 = 0;
    // ---- scaling and double-float coefficients
    double mx = 0;
#pragma unroll
    for (int k = 0; k <= DEG; ++k) mx = fmax(mx, fmax(fabs(p[k].re), fabs(p[k].im)));
    const double sc = (mx > 0) ? ldexp(1.0, -ilogb(mx)) : 1.0;
    cf ph[DEG + 1], pl[DEG + 1];
    float al[DEG + 1];
    double al64[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) {
        double re = p[k].re * sc, im = p[k].im * sc;
        ph[k] = cf{(float)re, (float)im};
        pl[k] = cf{(float)(re - (double)ph[k].re), (float)(im - (double)ph[k].im)};
        al64[k] = GUARD ? cabs_safe(p[k]) : cabs(p[k]);
        al[k] = (float)(al64[k] * sc) * (float)fma(ALPHA_W, (double)k, 1.0);
    }
    if (!custom_init) init_est<DEG, false>(al64, z);
#pragma unroll
    for (int k = 0; k <= DEG; ++k) al64[k] = al64[k] * fma(ALPHA_W, (double)k, 1.0);
    cf zf[DEG];
#pragma unroll
    for (int k = 0; k < DEG; ++k) zf[k] = cf{(float)z[k].re, (float)z[k].im};

    // ---- stage A: FP32 iterations
    constexpr unsigned ALL = (1u << DEG) - 1u;
    unsigned cx = 0;
    int it = 0;
    for (; it < itmax && cx != ALL; ++it) {
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            if (!((cx >> j) & 1u)) {
                ++st.it32;
                cf x = zf[j];
                float ax = sqrtf(norm2(x));
                float b = al[DEG];
#pragma unroll
                for (int k = DEG - 1; k >= 0; --k) b = fmaf(b, ax, al[k]);
                cf h = ph[DEG], hd = cf{0.f, 0.f};
#pragma unroll
                for (int k = DEG - 1; k >= 0; --k) {
                    hd = cfma(hd, x, h);
                    h = cfma(h, x, ph[k]);
                }
                const float tb = EA_EPS32 * b;
                if (norm2(h) > tb * tb) {
                    cf S = aberth_sum_f<DEG>(zf, j, x);
                    zf[j] = x - cdiv(h, hd - h * S);
                } else {
                    cx |= 1u << j;
                }
            }
        }
    }
    bool finite = true;
#pragma unroll
    for (int k = 0; k < DEG; ++k) finite = finite && isfinite(zf[k].re) && isfinite(zf[k].im);
    bool done = false;
    int it2 = 0;
    if (finite) {
    // ---- stage B: one compensated FP32 sweep, result in double-float
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
#pragma clang fp contract(off)
        ++st.itcomp;
        cf x = zf[j];
        cf h = ph[DEG], e = pl[DEG], hd = cf{0.f, 0.f};
#pragma unroll
        for (int k = DEG - 1; k >= 0; --k) {
            hd = cfma(hd, x, h);
            float p0, q0, p1, q1, p2, q2, p3, q3, s4, r4, s5, r5, sre, ere, sim, eim;
            two_prod_f(h.re, x.re, p0, q0);
            two_prod_f(h.im, x.im, p1, q1);
            two_prod_f(h.re, x.im, p2, q2);
            two_prod_f(h.im, x.re, p3, q3);
            two_sum_f(p0, -p1, s4, r4);
            two_sum_f(p2, p3, s5, r5);
            two_sum_f(s4, ph[k].re, sre, ere);
            two_sum_f(s5, ph[k].im, sim, eim);
            h = cf{sre, sim};
            cf err = cf{((q0 - q1) + r4) + (ere + pl[k].re), ((q2 + q3) + r5) + (eim + pl[k].im)};
            e = cfma(e, x, err);
        }
        cf hc = h + e;
        cf S = aberth_sum_f<DEG>(zf, j, x);
        cf corr = cdiv(hc, hd - hc * S);
        // x - corr in double-float, stored straight to FP64
        z[j] = cd{(double)x.re - (double)corr.re, (double)x.im - (double)corr.im};
        zf[j] = cf{(float)z[j].re, (float)z[j].im};
    }

    // ---- stage C: FP64 polish under the FP64 stopping test
    cx = 0;
    for (; it2 < itmax; ++it2) {
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            if (!((cx >> j) & 1u)) {
                ++st.it64;
                if (fast_step<DEG, GUARD>(p, al64, z, j)) cx |= 1u << j;
            }
        }
        if (cx == ALL) {
            done = true;
            ++it2;
            break;
        }
    }
    st.sweeps = it + 1 + it2;
#pragma unroll
    for (int k = 0; k < DEG; ++k) finite = finite && isfinite(z[k].re) && isfinite(z[k].im);
    }  // finite after stage A
    // A non-finite FP32 iterate, or two estimates that FP32 cannot tell apart (a nearly
    // double root: the Aberth sums of stages B and C turn infinite and NaN roots would
    // pass the stopping test): the polynomial is solved again by the FP64 form.
    if (!finite) {
        SolveStats s2;
        done = solve_fp64<DEG, MODE_FP64, false, GUARD>(p, z, itmax, false, s2);
        st.sweeps = it + s2.sweeps;
        st.it64 += s2.it64;
    }
    return done;
}

}  // namespace caus
