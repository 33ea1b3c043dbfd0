/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker (never as the thing measured on the GPU or shipped).
 *
 * Plain-C FP64 restatement of the reference's Ehrlich-Aberth all-roots solver
 * (CompEA port inside fbartolic/caustics):
 *   - init_est / convex_hull / ccw ........ lib/ehrlich_aberth/init_est.h:30-102
 *   - ehrlich_aberth (plain FP64) ......... lib/ehrlich_aberth/ehrlich_aberth.h:60-150
 *   - correction / rcorrection ............ lib/ehrlich_aberth/ehrlich_aberth.h:27-58
 *   - ehrlich_aberth_comp (EFT stage) ..... lib/ehrlich_aberth/ehrlich_aberth.h:153-292
 *   - horner_dble/rhorner_dble/horner_cmplx/rhorner_cmplx, EFT two_sum/two_prod,
 *     Priest summation, compensated Horner  lib/ehrlich_aberth/horner.h:21-385
 *   - batch loop (cpu_ehrlich_aberth) ..... lib/ehrlich_aberth/cpu_ops.cc:15-81
 * Complex arithmetic follows the vendored thrust::complex<double> semantics the
 * reference compiles against (division scaled by |re|+|im|, abs = hypot, pow(z,2)
 * = exp(2 log z)), lib/extern/thrust-1.15.0/thrust/detail/complex/arithmetic.h:119-160,
 * :184-189 and cpow.h:36-43.  The one documented difference: the complex log/exp
 * used for pow(z,2) are glibc's clog/cexp rather than thrust's FreeBSD-derived
 * ones (last-bit differences only).
 *
 * init_mode 0 reproduces the reference's init_est comma-operator defect
 * (init_est.h:95: every initial estimate is r*sin(theta) on the real axis, and
 * the reversed correction squares z via exp(2 log z)).  init_mode 1 is the
 * intended CompEA initialisation r*(cos theta + i sin theta) with z*z computed
 * exactly; both converge to the same root sets.
 *
 * Additions that do not change results: per-polynomial sweep counts, active
 * root-iteration counts split by branch (forward |z|<=1 / reversed |z|>1), and a
 * convergence status byte instead of printf+assert (ehrlich_aberth.h:144-149).
 * The batch entry runs independent polynomials on OpenMP threads.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static const double EPS = DBL_EPSILON / 2; /* horner.h:21 */

static inline cplx mk(double re, double im) { cplx z = {re, im}; return z; }
static inline cplx cadd(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
/* thrust operator*(complex, complex), arithmetic.h:88-95 */
static inline cplx cmul(cplx x, cplx y) {
    return mk(x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re);
}
static inline cplx cscal(double s, cplx y) { return mk(s * y.re, s * y.im); }
/* thrust operator/(complex, complex): scale by |re|+|im| first, arithmetic.h:119-142 */
static inline cplx cdiv(cplx x, cplx y) {
    double s = fabs(y.re) + fabs(y.im);
    double oos = 1.0 / s;
    double ars = x.re * oos, ais = x.im * oos;
    double brs = y.re * oos, bis = y.im * oos;
    s = (brs * brs) + (bis * bis);
    oos = 1.0 / s;
    return mk(((ars * brs) + (ais * bis)) * oos, ((ais * brs) - (ars * bis)) * oos);
}
static inline cplx crecip(cplx y) { return cdiv(mk(1.0, 0.0), y); }
static inline double cabs_(cplx z) { return hypot(z.re, z.im); } /* arithmetic.h:184-189 */
/* thrust::pow(z, 2) = exp(log(z) * 2.0), cpow.h:36-43 */
static inline cplx cpow2_explog(cplx z) {
    double complex zz = z.re + I * z.im;
    double complex r = cexp(clog(zz) * 2.0);
    return mk(creal(r), cimag(r));
}
static inline cplx csq(cplx z, int init_mode) {
    return init_mode == 0 ? cpow2_explog(z) : cmul(z, z);
}

/* ---------------- initial estimates: init_est.h:21-102 ---------------- */
typedef struct { int x; double y; } hpoint;

static double ccw(const hpoint *p1, const hpoint *p2, const hpoint *p3) {
    return (p2->x - p1->x) * (p3->y - p1->y) - (p2->y - p1->y) * (p3->x - p1->x);
}

static void convex_hull(const hpoint *points, int npoints, hpoint *hull, int *hullsize) {
    int k = 0;
    for (int i = npoints - 1; i >= 0; i--) {
        while (k >= 2 && ccw(&hull[k - 2], &hull[k - 1], &points[i]) <= 0) --k;
        hull[k++] = points[i];
    }
    *hullsize = k;
}

void ora_init_est(const double *alpha, int deg, cplx *roots, int init_mode) {
    hpoint points[32], hull[32];
    int hullsize;
    const double pi2 = 6.28318530717958647693, sigma = 0.7;
    double th = pi2 / deg;
    int k = 0;
    for (int i = 0; i <= deg; i++) {
        points[i].x = i;
        points[i].y = alpha[i] > 0 ? log(alpha[i]) : -1E+30;
    }
    convex_hull(points, deg + 1, hull, &hullsize);
    for (int i = hullsize - 2; i >= 0; i--) {
        int nzeros = hull[i].x - hull[i + 1].x;
        double a1 = pow(alpha[hull[i + 1].x], 1.0 / nzeros);
        double a2 = pow(alpha[hull[i].x], 1.0 / nzeros);
        double r = a1 / a2;
        double ang = pi2 / nzeros;
        for (int j = 0; j < nzeros; j++) {
            double t = ang * j + th * i + sigma;
            if (init_mode == 0)
                roots[k + j] = mk(r * sin(t), 0.0); /* init_est.h:95 as compiled */
            else
                roots[k + j] = mk(r * cos(t), r * sin(t));
        }
        k += nzeros;
    }
}

/* ---------------- Horner family: horner.h:219-267 ---------------- */
static void horner_dble(const double *p, double x, int deg, double *h) {
    *h = p[deg];
    for (int i = deg - 1; i >= 0; --i) *h = fma(*h, x, p[i]);
}
static void rhorner_dble(const double *p, double x, int deg, double *h) {
    *h = p[0];
    for (int i = 1; i <= deg; ++i) *h = fma(*h, x, p[i]);
}
static void horner_cmplx(const cplx *p, cplx x, int deg, cplx *h, cplx *hd) {
    *h = p[deg];
    *hd = mk(0, 0);
    for (int i = deg - 1; i >= 0; --i) {
        *hd = cadd(cmul(*hd, x), *h);
        *h = cadd(cmul(*h, x), p[i]);
    }
}
static void rhorner_cmplx(const cplx *p, cplx x, int deg, cplx *h, cplx *hd) {
    *h = p[0];
    *hd = mk(0, 0);
    for (int i = 1; i <= deg; ++i) {
        *hd = cadd(cmul(*hd, x), *h);
        *h = cadd(cmul(*h, x), p[i]);
    }
}

/* ---------------- EFTs: horner.h:30-84, 164-217 ---------------- */
static double gamma_const(unsigned n) {
    double s = 1.41421356237309504880;
    double g = (2 * n * EPS) * s;
    return g / ((1 - DBL_EPSILON) - g);
}
static inline void two_sum(double a, double b, double *s, double *e) {
    *s = a + b;
    double t = *s - a;
    *e = (a - (*s - t)) + (b - t);
}
static inline void two_prod(double a, double b, double *p, double *e) {
    *p = a * b;
    *e = fma(a, b, -*p);
}
/* result, and three error terms of a complex product (horner.h:69-84) */
static inline void two_prod_cmplx(cplx a, cplx b, cplx *res, cplx *e1, cplx *e2, cplx *e3) {
    double p0, q0, p1, q1, p2, q2, p3, q3, s4, r4, s5, r5;
    two_prod(a.re, b.re, &p0, &q0);
    two_prod(a.im, b.im, &p1, &q1);
    two_prod(a.re, b.im, &p2, &q2);
    two_prod(a.im, b.re, &p3, &q3);
    two_sum(p0, -p1, &s4, &r4);
    two_sum(p2, p3, &s5, &r5);
    *res = mk(s4, s5);
    *e1 = mk(q0, q2);
    *e2 = mk(-q1, q3);
    *e3 = mk(r4, r5);
}
static inline void two_sum_cmplx(cplx a, cplx b, cplx *res, cplx *err) {
    double s0, e0, s1, e1;
    two_sum(a.re, b.re, &s0, &e0);
    two_sum(a.im, b.im, &s1, &e1);
    *res = mk(s0, s1);
    *err = mk(e0, e1);
}
/* selection sort by decreasing magnitude (horner.h:164-188) */
static void sort4(double *p) {
    for (int i = 0; i < 3; ++i) {
        double mx = fabs(p[i]);
        int ind = i;
        for (int j = i + 1; j < 4; ++j) {
            double t = fabs(p[j]);
            if (t > mx) { mx = t; ind = j; }
        }
        if (ind != i) { double t = p[i]; p[i] = p[ind]; p[ind] = t; }
    }
}
/* Priest's doubly compensated summation of 4 terms (horner.h:190-208) */
static double priest_sum(double *p) {
    sort4(p);
    double s = p[0], c = 0, y, u, t, v, z;
    for (int i = 1; i < 4; ++i) {
        y = c + p[i];
        u = p[i] - (y - c);
        t = y + s;
        v = y - (t - s);
        z = u + v;
        s = t + z;
        c = z - (s - t);
    }
    return s;
}
static cplx priest_cmplx_sum(const cplx *p) {
    double re[4] = {p[0].re, p[1].re, p[2].re, p[3].re};
    double im[4] = {p[0].im, p[1].im, p[2].im, p[3].im};
    return mk(priest_sum(re), priest_sum(im));
}
/* compensated complex Horner with running error bound, horner.h:281-385.
 * reversed=0 evaluates p at x from the top coefficient; reversed=1 evaluates the
 * reversed polynomial (starting at p[0]). */
static void horner_comp_cmplx(const cplx *poly, cplx x, int deg, int reversed, cplx *h, cplx *hd,
                              double *eb) {
    cplx e = mk(0, 0), ed = mk(0, 0), p[4], res, e1, e2, e3, serr;
    double ap[4];
    *h = reversed ? poly[0] : poly[deg];
    *hd = mk(0, 0);
    *eb = 0;
    for (int step = 0; step < deg; ++step) {
        int i = reversed ? step + 1 : deg - 1 - step;
        two_prod_cmplx(*hd, x, &res, &e1, &e2, &e3);
        two_sum_cmplx(res, *h, hd, &serr);
        p[0] = e1; p[1] = e2; p[2] = e3; p[3] = serr;
        ed = cadd(cadd(cmul(ed, x), e), priest_cmplx_sum(p));
        two_prod_cmplx(*h, x, &res, &e1, &e2, &e3);
        two_sum_cmplx(res, poly[i], h, &serr);
        p[0] = e1; p[1] = e2; p[2] = e3; p[3] = serr;
        e = cadd(cmul(e, x), priest_cmplx_sum(p));
        ap[0] = cabs_(e1); ap[1] = cabs_(e2); ap[2] = cabs_(e3); ap[3] = cabs_(serr);
        *eb = *eb * cabs_(x) + priest_sum(ap);
    }
    *h = cadd(*h, e);
    *hd = cadd(*hd, ed);
}

/* ---------------- corrections: ehrlich_aberth.h:27-58 ---------------- */
static cplx aberth_sum(const cplx *roots, int deg, int j) {
    cplx corr = mk(0, 0);
    for (int i = 0; i < j; i++) corr = cadd(corr, crecip(csub(roots[j], roots[i])));
    for (int i = j + 1; i < deg; i++) corr = cadd(corr, crecip(csub(roots[j], roots[i])));
    return corr;
}
static cplx correction(const cplx *roots, cplx h, cplx hd, int deg, int j) {
    cplx corr = aberth_sum(roots, deg, j);
    return cdiv(h, csub(hd, cmul(h, corr)));
}
static cplx rcorrection(const cplx *roots, cplx h, cplx hd, int deg, int j, int init_mode) {
    cplx corr = aberth_sum(roots, deg, j);
    cplx z2 = csq(roots[j], init_mode);
    cplx num = cmul(z2, h);
    cplx den = csub(csub(cmul(cscal((double)deg, roots[j]), h), hd), cmul(cmul(z2, h), corr));
    return cdiv(num, den);
}

/* ---------------- solver: ehrlich_aberth.h:60-292 ----------------
 * Returns 1 when every root converged (the reference's `s`), else 0.
 * stats (nullable): [0]=sweeps, [1]=forward active root-iterations,
 * [2]=reversed active root-iterations, [3]=compensated-stage root-iterations. */
int ora_ehrlich_aberth(int deg, int itmax, int compensated, int custom_init, int init_mode,
                       const cplx *poly, const cplx *roots_init, cplx *roots, int64_t *stats) {
    double alpha[32];
    int convx[32], convy[32];
    int s = 0, sweeps = 0;
    int64_t nf = 0, nr = 0, nc = 0;
    for (int i = 0; i < deg; i++) {
        alpha[i] = cabs_(poly[i]);
        convx[i] = 0;
        convy[i] = 0;
    }
    alpha[deg] = cabs_(poly[deg]);
    if (!custom_init)
        ora_init_est(alpha, deg, roots, init_mode);
    else
        for (int i = 0; i < deg; i++) roots[i] = roots_init[i];
    for (int i = 0; i <= deg; i++) alpha[i] = alpha[i] * fma(3.8284271247461900976, i, 1);

    for (int it = 0; it < itmax; it++) {
        sweeps = it + 1;
        for (int j = 0; j < deg; j++) {
            double b;
            cplx h, hd;
            if (!convx[j]) {
                if (cabs_(roots[j]) > 1) {
                    nr++;
                    rhorner_dble(alpha, 1. / cabs_(roots[j]), deg, &b);
                    rhorner_cmplx(poly, crecip(roots[j]), deg, &h, &hd);
                    if (cabs_(h) > EPS * b)
                        roots[j] = csub(roots[j], rcorrection(roots, h, hd, deg, j, init_mode));
                    else
                        convx[j] = 1;
                } else {
                    nf++;
                    horner_dble(alpha, cabs_(roots[j]), deg, &b);
                    horner_cmplx(poly, roots[j], deg, &h, &hd);
                    if (cabs_(h) > EPS * b)
                        roots[j] = csub(roots[j], correction(roots, h, hd, deg, j));
                    else
                        convx[j] = 1;
                }
            } else if (compensated && !convy[j]) {
                nc++;
                int rev = cabs_(roots[j]) > 1;
                cplx x = rev ? crecip(roots[j]) : roots[j];
                horner_comp_cmplx(poly, x, deg, rev, &h, &hd, &b);
                double errBound =
                    EPS * cabs_(h) + (gamma_const(4 * deg + 2) * b + 2 * pow(EPS, 2) * cabs_(h));
                if (cabs_(h) > 4 * errBound) {
                    cplx corr = rev ? rcorrection(roots, h, hd, deg, j, init_mode)
                                    : correction(roots, h, hd, deg, j);
                    double thr = rev ? 4 * EPS * cabs_(roots[j]) : 4 * EPS;
                    if (cabs_(corr) > thr)
                        roots[j] = csub(roots[j], corr);
                    else
                        convy[j] = 1;
                } else {
                    convy[j] = 1;
                }
            }
        }
        s = 1;
        for (int j = 0; j < deg; j++) s = s && (compensated ? convy[j] : convx[j]);
        if (s) break;
    }
    if (stats) {
        stats[0] = sweeps;
        stats[1] = nf;
        stats[2] = nr;
        stats[3] = nc;
    }
    return s;
}

/* Batch entry mirroring cpu_ehrlich_aberth (cpu_ops.cc:15-81): coeffs (N, deg+1)
 * ascending, interleaved re/im; roots (N, deg).  status[i] = 1 when polynomial i
 * converged.  nthreads <= 0 uses the OpenMP default.  Returns the number of
 * unconverged polynomials. */
int64_t ora_roots_batch(int64_t N, int deg, int itmax, int compensated, int custom_init,
                        int init_mode, const double *coeffs, const double *roots_init,
                        double *roots, int64_t *stats, uint8_t *status, int nthreads) {
    int64_t bad = 0;
    if (deg < 1 || deg > 30) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
#endif
    for (int64_t i = 0; i < N; i++) {
        const cplx *p = (const cplx *)(coeffs + 2 * i * (deg + 1));
        const cplx *ri = roots_init ? (const cplx *)(roots_init + 2 * i * deg) : NULL;
        cplx *r = (cplx *)(roots + 2 * i * deg);
        int ok = ora_ehrlich_aberth(deg, itmax, compensated, custom_init && ri, init_mode, p, ri,
                                    r, stats ? stats + 4 * i : NULL);
        if (status) status[i] = (uint8_t)ok;
        bad += !ok;
    }
    return bad;
}

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------- canonical polish (no counterpart in the reference) ----------------
 * The same canonicalisation the product applies to every contour limb solve: Ehrlich-
 * Aberth sweeps whose residual and derivative come from compensated (double-double)
 * Horner, run while some root still moves by more than 2^-40 of its modulus (at most 8
 * sweeps; the first sweep takes a plain derivative).  Converged roots end at the doubles
 * nearest the exact roots of the given FP64 polynomial, which makes image decisions taken
 * on them (the 1e-6 lens-equation filter) reproducible where the reference's backward-
 * error stopping test leaves ill-conditioned roots far from them.  Written from the
 * method (Graillat-Langlois-Louvet compensated Horner), not from the product's code. */
static void horner_comp_vd(const cplx *p, cplx x, int deg, int deriv_comp, cplx *h, cplx *hd) {
    cplx e = mk(0, 0), ed = mk(0, 0), res, e1, e2, e3, se;
    *h = p[deg];
    *hd = mk(0, 0);
    for (int i = deg - 1; i >= 0; --i) {
        if (deriv_comp) {
            two_prod_cmplx(*hd, x, &res, &e1, &e2, &e3);
            two_sum_cmplx(res, *h, hd, &se);
            ed = cadd(cadd(cmul(ed, x), e), mk((e1.re + e2.re) + (e3.re + se.re),
                                                (e1.im + e2.im) + (e3.im + se.im)));
        } else {
            *hd = cadd(cmul(*hd, x), *h);
        }
        two_prod_cmplx(*h, x, &res, &e1, &e2, &e3);
        two_sum_cmplx(res, p[i], h, &se);
        e = cadd(cmul(e, x), mk((e1.re + e2.re) + (e3.re + se.re), (e1.im + e2.im) + (e3.im + se.im)));
    }
    *h = cadd(*h, e);
    *hd = cadd(*hd, ed);
}

void ora_polish(int deg, const cplx *p, cplx *z) {
    unsigned act = (deg >= 32) ? 0xFFFFFFFFu : ((1u << deg) - 1u);
    for (int sweep = 0; sweep < 8 && act; ++sweep) {
        unsigned nxt = 0;
        for (int j = 0; j < deg; ++j) {
            if (!((act >> j) & 1u)) continue;
            cplx h, hd;
            horner_comp_vd(p, z[j], deg, sweep > 0, &h, &hd);
            cplx den = csub(hd, cmul(h, aberth_sum(z, deg, j)));
            if (den.re == 0 && den.im == 0) continue;
            cplx corr = cdiv(h, den);
            cplx zn = csub(z[j], corr);
            if (!isfinite(zn.re) || !isfinite(zn.im)) continue;
            double nc = corr.re * corr.re + corr.im * corr.im;
            double nz = z[j].re * z[j].re + z[j].im * z[j].im;
            if (nc > 0x1p-80 * nz) nxt |= 1u << j;
            z[j] = zn;
        }
        act = nxt;
    }
}

void ora_polish_batch(int64_t N, int deg, const double *coeffs, double *roots, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < N; i++)
        ora_polish(deg, (const cplx *)(coeffs + 2 * i * (deg + 1)), (cplx *)(roots + 2 * i * deg));
}
