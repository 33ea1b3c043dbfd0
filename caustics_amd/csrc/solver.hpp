// Ehrlich-Aberth all-roots solver, one polynomial per lane, for gfx950.
//
// Algorithm (reference, fbartolic/caustics):
//   initial estimates ........ lib/ehrlich_aberth/init_est.h:57-102 (upper convex hull of
//                              (i, log|p_i|), one ring of roots per hull edge)
//   FP64 iteration ........... lib/ehrlich_aberth/ehrlich_aberth.h:60-150
//   compensated stage ........ lib/ehrlich_aberth/ehrlich_aberth.h:153-292,
//                              lib/ehrlich_aberth/horner.h:281-385
// Layout: coefficients ascending (p[k] multiplies z^k), held in VGPRs; every
// loop over the degree is unrolled at compile time, so the iteration's arrays stay in
// registers (the RefCompat form's init_est reads its log-magnitude table at runtime hull
// vertices, the one array the compiler places in scratch).
//
// Two iteration forms:
//   RefCompat  the reference's branch structure: |z|>1 uses the reversed polynomial
//              at 1/z and squares z as exp(2 log z) (ehrlich_aberth.h:56-57); with
//              real-axis initial estimates (init_est.h:95 as compiled) it reproduces
//              the reference's root order.
//   Fast       complex initial estimates r(cos t + i sin t); every root uses the
//              forward polynomial.  The forward and reversed Aberth steps are the same
//              rational function of z and their stopping tests |h| <= eps*b compare the
//              same ratio (both sides scale by |z|^-n), so only rounding differs; with
//              degree <= 12 and |z| bounded by the Fujiwara radius there is no overflow.
#pragma once
#include "cplx.hpp"

namespace caus {

constexpr double EA_EPS = 1.1102230246251565e-16;  // DBL_EPSILON/2, horner.h:21
constexpr double ALPHA_W = 3.8284271247461900976;   // 1+2*sqrt(2), ehrlich_aberth.h:96-99

enum SolveMode : int { MODE_FP64 = 0, MODE_FP64_COMP = 1, MODE_CF32 = 2 };

// per-polynomial counters: [0] sweeps, [1] FP64 root-iterations (plain stage),
// [2] compensated / double-float root-iterations, [3] FP32 root-iterations
struct SolveStats {
    int sweeps, it64, itcomp, it32;
};

template <int N>
__device__ __forceinline__ double sel(const double (&a)[N], int idx) {
    double v = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) v = (idx == k) ? a[k] : v;
    return v;
}

// init_est.h:57-102.  The hull is kept as a bitmask of abscissae; points are pushed
// from x = DEG down to 0 exactly as convex_hull (init_est.h:37-54) does, so the top of
// the stack is the lowest set bit.  Root slot m lies on the hull edge [xlo, xhi)
// containing m, which is where the reference's running offset k places it.
// Fast form: the same hull found with compile-time indices only, so that the table
// stays in registers.  Point i (0 < i < DEG) is a vertex of the upper hull exactly when
// a line through it passes strictly above every other point, i.e. when every slope
// from a point on its left exceeds every slope to a point on its right; the end points
// are always vertices.  (The monotone chain pops collinear points, so equality is no
// vertex in either test; the two tests can differ only by the rounding of nearly
// collinear triples, which moves an initial estimate, not a root.)  Root slot m then
// lies on the edge [xlo(m), xhi(m)) of the nearest vertices at or below and above m.
template <int DEG>
__device__ __forceinline__ void init_est_static(const double (&y)[DEG + 1], cd (&z)[DEG]) {
    bool vx[DEG + 1];
    vx[0] = vx[DEG] = true;
#pragma unroll
    for (int i = 1; i < DEG; ++i) {
        double lo = (y[i] - y[0]) * (1.0 / i), hi = (y[DEG] - y[i]) * (1.0 / (DEG - i));
#pragma unroll
        for (int a = 1; a < i; ++a) lo = fmin(lo, (y[i] - y[a]) * (1.0 / (i - a)));
#pragma unroll
        for (int b = i + 1; b < DEG; ++b) hi = fmax(hi, (y[b] - y[i]) * (1.0 / (b - i)));
        vx[i] = hi < lo;
    }
    // nearest vertex above each slot, and the number of vertices above it
    int xhi[DEG], nabove[DEG];
    double yhi[DEG];
    {
        int x = DEG, na = 1;   // DEG is a vertex
        double yy = y[DEG];
#pragma unroll
        for (int m = DEG - 1; m >= 0; --m) {
            xhi[m] = x;
            yhi[m] = yy;
            nabove[m] = na;
            if (vx[m]) {
                ++na;
                x = m;
                yy = y[m];
            }
        }
    }
    const double sigma_pi = 0.7 * 0.31830988618379067154;
    int xlo = 0;
    double ylo = y[0];
#pragma unroll
    for (int m = 0; m < DEG; ++m) {
        if (vx[m]) {
            xlo = m;
            ylo = y[m];
        }
        const int nz = xhi[m] - xlo;
        const double r = exp((ylo - yhi[m]) / nz);
        // e = vertices strictly above xhi: nabove[m] counts those above m, xhi included
        const int e = nabove[m] - 1;
        double s, c;
        sincospi((2.0 / nz) * (m - xlo) + (2.0 / DEG) * e + sigma_pi, &s, &c);
        z[m] = cd{r * c, r * s};
    }
}

template <int DEG, bool REALAXIS>
__device__ __forceinline__ void init_est(const double (&alpha)[DEG + 1], cd (&z)[DEG]) {
    double y[DEG + 1];
#pragma unroll
    for (int i = 0; i <= DEG; ++i) y[i] = alpha[i] > 0 ? log(alpha[i]) : -1e30;
    if constexpr (!REALAXIS) {
        init_est_static<DEG>(y, z);
        return;
    }
    unsigned st = 0;
#pragma unroll
    for (int i = DEG; i >= 0; --i) {
        while (__builtin_popcount(st) >= 2) {
            int x2 = __builtin_ctz(st);
            int x1 = __builtin_ctz(st & (st - 1));
            double y1 = sel(y, x1), y2 = sel(y, x2);
            double c = (double)(x2 - x1) * (y[i] - y1) - (y2 - y1) * (double)(i - x1);
            if (c <= 0)
                st &= st - 1;
            else
                break;
        }
        st |= 1u << i;
    }
    const double pi2 = 6.28318530717958647693, sigma = 0.7, th = pi2 / DEG;
#pragma unroll
    for (int m = 0; m < DEG; ++m) {
        const unsigned low = (2u << m) - 1u;
        int xlo = 31 - __builtin_clz(st & low);
        int xhi = __builtin_ctz(st & ~low);
        int nz = xhi - xlo;
        int j = m - xlo;
        int e = __builtin_popcount(st >> (xhi + 1));
        double r = exp((sel(y, xlo) - sel(y, xhi)) / nz);
        if (REALAXIS) {
            const double t = (pi2 / nz) * j + th * e + sigma;
            z[m] = cd{r * sin(t), 0.0};
        } else {
            // the same angle in units of pi: sincospi needs no long argument reduction,
            // which keeps this code small where it is inlined into the sweep kernel
            double s, c;
            sincospi((2.0 / nz) * j + (2.0 / DEG) * e + sigma * 0.31830988618379067154, &s, &c);
            z[m] = cd{r * c, r * s};
        }
    }
}

// exp(2 log z) (thrust::pow(z, 2), cpow.h:36-43)
__device__ __forceinline__ cd pow2_explog(cd z) {
    double lr = log(sqrt(norm2(z)));
    double th = atan2(z.im, z.re);
    double m = exp(2.0 * lr), s, c;
    sincos(2.0 * th, &s, &c);
    return cd{m * c, m * s};
}

// sum_{i != j} 1/(x - z_i).  RCP_STEPS Newton steps refine the v_rcp_f64 seed (relative
// error 2^-24 unrefined): the Fast form uses the bare seed, since an Aberth correction
// with a slightly perturbed sum is still a convergent iteration and the accuracy of the
// result is set by the exact FP64 residual in the stopping test.
template <int DEG, int RCP_STEPS = 1>
__device__ __forceinline__ cd aberth_sum(const cd (&z)[DEG], int j, cd x) {
    cd S = cd{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < DEG; ++i)
        if (i != j) S = S + crcp<RCP_STEPS>(x - z[i]);
    return S;
}

// ---- EFT building blocks (horner.h:37-84, 164-217) ---------------------------
// Error-free transformations must round every operation separately: FMA contraction
// across these statements (a product feeding a later sum) would break the error terms.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    double t = s - a;
    e = (a - (s - t)) + (b - t);
}
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
#pragma clang fp contract(off)
    p = a * b;
    e = fma(a, b, -p);
}
__device__ __forceinline__ void two_prod_c(cd a, cd b, cd& r, cd& e1, cd& e2, cd& e3) {
#pragma clang fp contract(off)
    double p0, q0, p1, q1, p2, q2, p3, q3, s4, r4, s5, r5;
    two_prod(a.re, b.re, p0, q0);
    two_prod(a.im, b.im, p1, q1);
    two_prod(a.re, b.im, p2, q2);
    two_prod(a.im, b.re, p3, q3);
    two_sum(p0, -p1, s4, r4);
    two_sum(p2, p3, s5, r5);
    r = cd{s4, s5};
    e1 = cd{q0, q2};
    e2 = cd{-q1, q3};
    e3 = cd{r4, r5};
}
__device__ __forceinline__ void two_sum_c(cd a, cd b, cd& s, cd& e) {
    two_sum(a.re, b.re, s.re, e.re);
    two_sum(a.im, b.im, s.im, e.im);
}
__device__ __forceinline__ void cswap_desc(double& a, double& b) {
    bool sw = fabs(b) > fabs(a);
    double ta = sw ? b : a, tb = sw ? a : b;
    a = ta;
    b = tb;
}
// Priest's doubly compensated sum of four terms sorted by decreasing magnitude
// (horner.h:164-208); the sort is a 5-comparator network.
__device__ __forceinline__ double priest4(double p0, double p1, double p2, double p3) {
#pragma clang fp contract(off)
    cswap_desc(p0, p1);
    cswap_desc(p2, p3);
    cswap_desc(p0, p2);
    cswap_desc(p1, p3);
    cswap_desc(p1, p2);
    double s = p0, c = 0, y, u, t, v, zz;
    double q[3] = {p1, p2, p3};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        y = c + q[i];
        u = q[i] - (y - c);
        t = y + s;
        v = y - (t - s);
        zz = u + v;
        s = t + zz;
        c = zz - (s - t);
    }
    return s;
}
__device__ __forceinline__ double gamma_const(int n) {
    double g = (2 * n * EA_EPS) * 1.41421356237309504880;
    return g / ((1 - 2 * EA_EPS) - g);
}

// compensated complex Horner with running error bound (horner.h:281-385)
template <int DEG, bool REV>
__device__ __forceinline__ void horner_comp(const cd (&p)[DEG + 1], cd x, cd& h, cd& hd,
                                            double& eb) {
#pragma clang fp contract(off)
    cd e = cd{0, 0}, ed = cd{0, 0};
    h = REV ? p[0] : p[DEG];
    hd = cd{0, 0};
    eb = 0;
    double ax = cabs(x);
#pragma unroll
    for (int step = 0; step < DEG; ++step) {
        const int i = REV ? step + 1 : DEG - 1 - step;
        cd r, e1, e2, e3, se;
        two_prod_c(hd, x, r, e1, e2, e3);
        two_sum_c(r, h, hd, se);
        ed = ed * x + e +
             cd{priest4(e1.re, e2.re, e3.re, se.re), priest4(e1.im, e2.im, e3.im, se.im)};
        two_prod_c(h, x, r, e1, e2, e3);
        two_sum_c(r, p[i], h, se);
        e = e * x + cd{priest4(e1.re, e2.re, e3.re, se.re), priest4(e1.im, e2.im, e3.im, se.im)};
        eb = eb * ax + priest4(cabs(e1), cabs(e2), cabs(e3), cabs(se));
    }
    h = h + e;
    hd = hd + ed;
}

// Value of the polynomial by compensated Horner (TwoProd/TwoSum error terms summed in
// plain FP64, accurate as if computed in twice the working precision) and its derivative
// by plain Horner.
template <int DEG>
__device__ __forceinline__ void horner_comp_value(const cd (&p)[DEG + 1], cd x, cd& h, cd& hd) {
#pragma clang fp contract(off)
    cd e = cd{0, 0};
    h = p[DEG];
    hd = cd{0, 0};
#pragma unroll
    for (int i = DEG - 1; i >= 0; --i) {
        hd = cd{fma(hd.re, x.re, fma(-hd.im, x.im, h.re)), fma(hd.re, x.im, fma(hd.im, x.re, h.im))};
        cd r, e1, e2, e3, se;
        two_prod_c(h, x, r, e1, e2, e3);
        two_sum_c(r, p[i], h, se);
        const cd t = cd{(e1.re + e2.re) + (e3.re + se.re), (e1.im + e2.im) + (e3.im + se.im)};
        e = cd{fma(e.re, x.re, fma(-e.im, x.im, t.re)), fma(e.re, x.im, fma(e.im, x.re, t.im))};
    }
    h = h + e;
}

// Value and derivative of the polynomial, both by compensated Horner: the derivative
// recurrence hd <- hd x + h carries its own TwoProd/TwoSum error terms, plus the error
// of the value it consumes.  At a nearly multiple root p'(z) is far smaller than the
// rounding error of a plain FP64 derivative, which would leave a Newton or Aberth
// correction there without a single correct digit.
template <int DEG>
__device__ __forceinline__ void horner_comp_vd(const cd (&p)[DEG + 1], cd x, cd& h, cd& hd) {
#pragma clang fp contract(off)
    cd e = cd{0, 0}, ed = cd{0, 0};
    h = p[DEG];
    hd = cd{0, 0};
#pragma unroll
    for (int i = DEG - 1; i >= 0; --i) {
        cd r, e1, e2, e3, se;
        two_prod_c(hd, x, r, e1, e2, e3);
        two_sum_c(r, h, hd, se);
        const cd td = cd{(e1.re + e2.re) + (e3.re + se.re), (e1.im + e2.im) + (e3.im + se.im)};
        ed = cd{fma(ed.re, x.re, fma(-ed.im, x.im, e.re + td.re)),
                fma(ed.re, x.im, fma(ed.im, x.re, e.im + td.im))};
        two_prod_c(h, x, r, e1, e2, e3);
        two_sum_c(r, p[i], h, se);
        const cd t = cd{(e1.re + e2.re) + (e3.re + se.re), (e1.im + e2.im) + (e3.im + se.im)};
        e = cd{fma(e.re, x.re, fma(-e.im, x.im, t.re)), fma(e.re, x.im, fma(e.im, x.re, t.im))};
    }
    h = h + e;
    hd = hd + ed;
}

// Canonical roots: Ehrlich-Aberth sweeps with the compensated value and derivative
// (accurate as if computed in twice the working precision; Aberth sum in FP64).  A root
// that met the FP64 stopping test lies within a few ulps of an exact root of the FP64
// polynomial, so one such step takes it to a small fraction of an ulp from it: the
// result is the double nearest the exact root whatever iteration converged to it (FP64
// or CF32, other initial estimates or warm starts).  Sweeps repeat while some root
// still moved by more than 2^-40 of its modulus, which happens only inside clusters of
// nearly coincident roots (images at a caustic, images next to a lens of very small
// mass), where the FP64 stopping test accepts iterates far from the roots; there the
// Aberth term keeps the cluster's roots apart and the compensated derivative keeps the
// corrections accurate, so they converge to the nearest doubles as well.  The discrete
// decisions taken downstream (image filter, continuation, segments, stitching) then see
// the same doubles in both precision forms.
// The first sweep takes the derivative in plain FP64: away from clusters its relative
// error only scales a correction of a few ulps, and most roots need no second sweep.  A
// root whose correction was below 2^-40 of its modulus is at its nearest double after
// that step (its error was squared), so later sweeps visit only the roots that moved more.
constexpr int kPolishSweeps = 8;
template <int DEG>
__device__ __forceinline__ void polish_roots(const cd (&p)[DEG + 1], cd (&z)[DEG]) {
    unsigned act = (1u << DEG) - 1u;
    for (int sweep = 0; sweep < kPolishSweeps && act; ++sweep) {
        unsigned nxt = 0;
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            if ((act >> j) & 1u) {
                cd h, hd;
                if (sweep == 0)
                    horner_comp_value<DEG>(p, z[j], h, hd);
                else
                    horner_comp_vd<DEG>(p, z[j], h, hd);
                const cd den = hd - h * aberth_sum<DEG, 2>(z, j, z[j]);
                if (norm2(den) > 0.0) {
                    const cd corr = cdiv<2>(h, den);
                    const cd zn = z[j] - corr;
                    if (isfinite(zn.re) && isfinite(zn.im)) {
                        if (norm2(corr) > 0x1p-80 * norm2(z[j])) nxt |= 1u << j;
                        z[j] = zn;
                    }
                }
            }
        }
        act = nxt;
    }
}

// One compensated-stage step for root j (ehrlich_aberth.h:226-280).  Returns true
// when the root is declared converged at the compensated level.
template <int DEG, bool REFCOMPAT>
__device__ __forceinline__ bool comp_step(const cd (&p)[DEG + 1], cd (&z)[DEG], int j) {
    cd zj = z[j];
    double az = cabs(zj);
    cd h, hd;
    double b;
    const double g = gamma_const(4 * DEG + 2);
    if (az > 1) {
        cd x = crcp<2>(zj);
        horner_comp<DEG, true>(p, x, h, hd, b);
        double ah = cabs_safe(h);
        double eb = EA_EPS * ah + (g * b + 2 * EA_EPS * EA_EPS * ah);
        if (ah > 4 * eb) {
            cd S = aberth_sum<DEG>(z, j, zj);
            cd z2 = REFCOMPAT ? pow2_explog(zj) : zj * zj;
            cd num = z2 * h;
            cd den = ((double)DEG * zj) * h - hd - num * S;
            cd corr = cdiv<2>(num, den);
            if (cabs_safe(corr) > 4 * EA_EPS * az) {
                z[j] = zj - corr;
                return false;
            }
        }
        return true;
    } else {
        horner_comp<DEG, false>(p, zj, h, hd, b);
        double ah = cabs_safe(h);
        double eb = EA_EPS * ah + (g * b + 2 * EA_EPS * EA_EPS * ah);
        if (ah > 4 * eb) {
            cd S = aberth_sum<DEG>(z, j, zj);
            cd corr = cdiv<2>(h, hd - h * S);
            if (cabs_safe(corr) > 4 * EA_EPS) {
                z[j] = zj - corr;
                return false;
            }
        }
        return true;
    }
}

// Plain FP64 step for root j, forward polynomial only (Fast form).  |x| in the bound
// b = sum alpha'_k |x|^k of the stopping test comes from the bare v_sqrt_f64 (the bound
// scales a threshold; an error of a few ulps in it moves the test negligibly), the
// residual h is exact FP64 Horner.
// Step for root j of a polynomial whose stopping bound tb = eps*b leaves the range in
// which squares are normal numbers (roots far below or above 1 with the coefficients
// scaled to O(1)): the test |h| > tb and the correction h/(hd - h S) in power-of-two
// scaled arithmetic, as thrust::abs (hypot) and thrust's scaled division do it in the
// reference.  Returns true when the root is converged.
template <int DEG>
__device__ __forceinline__ bool scaled_step(cd (&z)[DEG], int j, cd h, cd hd, double tb) {
    const cd x = z[j];
    // tb == 0: a root at 0 exactly.  tb NaN: the root itself is NaN and no step can mend
    // it, so it stops iterating (as the plain test's false comparison stops it) and the
    // polynomial is reported unconverged by roots_finite.
    if (!(tb > 0.0)) return tb == 0.0 ? cmaxabs(h) == 0.0 : true;
    if (!isfinite(tb)) return false;   // |x| beyond the range of doubles: not converged
    const double s = pow2_inv(tb);
    const cd hs = cd{h.re * s, h.im * s};
    const double ts = tb * s;
    if (!(norm2(hs) > ts * ts)) return true;
    cd S = cd{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < DEG; ++i)
        if (i != j) S = S + crcp_scaled(x - z[i]);
    z[j] = x - cdiv_scaled(h, hd - h * S);
    return false;
}

// GUARD: route polynomials whose squared stopping bound is not a normal number to
// scaled_step (the public root solver; the lens pipelines' polynomials have O(1)
// coefficients and roots and skip the two comparisons).
template <int DEG, bool GUARD = false>
__device__ __forceinline__ bool fast_step(const cd (&p)[DEG + 1], const double (&al)[DEG + 1],
                                          cd (&z)[DEG], int j) {
    cd x = z[j];
    const double ax = __builtin_amdgcn_sqrt(norm2(x));
    double b = al[DEG];
#pragma unroll
    for (int k = DEG - 1; k >= 0; --k) b = fma(b, ax, al[k]);
    cd h = p[DEG], hd = cd{0.0, 0.0};
#pragma unroll
    for (int k = DEG - 1; k >= 0; --k) {
        hd = cfma(hd, x, h);
        h = cfma(h, x, p[k]);
    }
    const double tb = EA_EPS * b;
    if (GUARD && !(tb >= 0x1p-480 && tb <= 0x1p480)) {
        // |x| itself may have underflowed in norm2: recompute b with a scaled modulus
        const double axs = cabs_safe(x);
        double bs = al[DEG];
#pragma unroll
        for (int k = DEG - 1; k >= 0; --k) bs = fma(bs, axs, al[k]);
        return scaled_step<DEG>(z, j, h, hd, EA_EPS * bs);
    }
    if (norm2(h) > tb * tb) {
        cd S = aberth_sum<DEG, 0>(z, j, x);
        z[j] = x - cdiv<2>(h, hd - h * S);
        return false;
    }
    return true;
}

// Plain FP64 step for root j with the reference's branch structure (RefCompat form).
template <int DEG, bool GUARD = false>
__device__ __forceinline__ bool ref_step(const cd (&p)[DEG + 1], const double (&al)[DEG + 1],
                                         cd (&z)[DEG], int j, int& nrev) {
    cd zj = z[j];
    double az = GUARD ? cabs_safe(zj) : cabs(zj);
    double b;
    cd h, hd = cd{0.0, 0.0};
    if (az > 1) {
        ++nrev;
        double ax = rcp(az);
        cd x = GUARD ? crcp_scaled(zj) : crcp<2>(zj);
        b = al[0];
#pragma unroll
        for (int k = 1; k <= DEG; ++k) b = fma(b, ax, al[k]);
        h = p[0];
#pragma unroll
        for (int k = 1; k <= DEG; ++k) {
            hd = cfma(hd, x, h);
            h = cfma(h, x, p[k]);
        }
        if ((GUARD ? cabs_safe(h) : cabs(h)) > EA_EPS * b) {
            cd S = aberth_sum<DEG>(z, j, zj);
            cd z2 = pow2_explog(zj);
            cd num = z2 * h;
            cd den = ((double)DEG * zj) * h - hd - num * S;
            z[j] = zj - cdiv<2>(num, den);
            return false;
        }
        return true;
    } else {
        b = al[DEG];
#pragma unroll
        for (int k = DEG - 1; k >= 0; --k) b = fma(b, az, al[k]);
        h = p[DEG];
#pragma unroll
        for (int k = DEG - 1; k >= 0; --k) {
            hd = cfma(hd, zj, h);
            h = cfma(h, zj, p[k]);
        }
        if (GUARD && !(EA_EPS * b >= 0x1p-480)) return scaled_step<DEG>(z, j, h, hd, EA_EPS * b);
        if (cabs(h) > EA_EPS * b) {
            cd S = aberth_sum<DEG>(z, j, zj);
            z[j] = zj - cdiv<2>(h, hd - h * S);
            return false;
        }
        return true;
    }
}

// Iteration state of one polynomial (registers of one lane).  GUARD: see fast_step.
template <int DEG, int MODE, bool REFCOMPAT, bool GUARD = false>
struct EAState {
    double al[DEG + 1];  // alpha' weights of the stopping test
    unsigned cx, cy;     // per-root convergence: plain stage, compensated stage
    int nrev;

    // alpha_i = |p_i| (ehrlich_aberth.h:77-82), initial estimates unless supplied,
    // alpha_i *= 1 + (1+2 sqrt 2) i (:96-99)
    __device__ __forceinline__ void start(const cd (&p)[DEG + 1], cd (&z)[DEG], bool custom_init) {
#pragma unroll
        for (int i = 0; i <= DEG; ++i) al[i] = GUARD ? cabs_safe(p[i]) : cabs(p[i]);
        if (!custom_init) init_est<DEG, REFCOMPAT>(al, z);
#pragma unroll
        for (int i = 0; i <= DEG; ++i) al[i] = al[i] * fma(ALPHA_W, (double)i, 1.0);
        cx = cy = 0;
        nrev = 0;
    }
    // one Gauss-Seidel sweep over the roots (ehrlich_aberth.h:102-143); returns true
    // when every root has converged
    __device__ __forceinline__ bool sweep(const cd (&p)[DEG + 1], cd (&z)[DEG], SolveStats& st) {
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            if (!((cx >> j) & 1u)) {
                ++st.it64;
                bool c = REFCOMPAT ? ref_step<DEG, GUARD>(p, al, z, j, nrev)
                                   : fast_step<DEG, GUARD>(p, al, z, j);
                if (c) cx |= 1u << j;
            } else if (MODE == MODE_FP64_COMP && !((cy >> j) & 1u)) {
                ++st.itcomp;
                if (comp_step<DEG, REFCOMPAT>(p, z, j)) cy |= 1u << j;
            }
        }
        ++st.sweeps;
        constexpr unsigned ALL = (1u << DEG) - 1u;
        return (MODE == MODE_FP64_COMP) ? (cy == ALL) : (cx == ALL);
    }
};

// Every root finite: a NaN or infinite root passes the stopping test (its comparisons
// are false) but is no solution, so a solve counts as converged only with this.
template <int DEG>
__device__ __forceinline__ bool roots_finite(const cd (&z)[DEG]) {
    bool f = true;
#pragma unroll
    for (int k = 0; k < DEG; ++k) f = f && isfinite(z[k].re) && isfinite(z[k].im);
    return f;
}

// Solve one polynomial held in registers.  Returns true when all roots converged
// (at the compensated level for MODE_FP64_COMP) to finite values.  z holds the
// initial estimates on entry when custom_init is set.
template <int DEG, int MODE, bool REFCOMPAT, bool GUARD = false>
__device__ __forceinline__ bool solve_fp64(const cd (&p)[DEG + 1], cd (&z)[DEG], int itmax,
                                           bool custom_init, SolveStats& st) {
    EAState<DEG, MODE, REFCOMPAT, GUARD> S;
    S.start(p, z, custom_init);
    st.sweeps = 0;
    st.it64 = 0;
    st.itcomp = 0;
    st.it32 = 0;
    bool done = false;
    for (int it = 0; it < itmax && !done; ++it) done = S.sweep(p, z, st);
    return done && roots_finite<DEG>(z);
}

// Coefficients whose largest component lies outside [2^-256, 2^256] are scaled by a
// power of two into [1, 2): exact, the roots are unchanged, and the squares of the
// solver's moduli stay normal numbers (the reference's hypot-based |.| and scaled
// division never square).  Other polynomials are left bit for bit as given.
template <int DEG>
__device__ __forceinline__ void prescale_coeffs(cd (&p)[DEG + 1]) {
    double m = 0.0;
#pragma unroll
    for (int k = 0; k <= DEG; ++k) m = fmax(m, cmaxabs(p[k]));
    if (m > 0.0 && isfinite(m) && !(m >= 0x1p-256 && m <= 0x1p256)) {
        const double s = pow2_inv(m);
#pragma unroll
        for (int k = 0; k <= DEG; ++k) p[k] = cd{p[k].re * s, p[k].im * s};
    }
}

}  // namespace caus
