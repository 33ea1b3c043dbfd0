/* C FP64 restatement of the point-source magnification, with OpenMP over source
 * positions.  TEST INFRASTRUCTURE ONLY: the host CPU baseline of the point-source and
 * light-curve workloads, and a cross-check of oracle/lens.py, which it follows step by
 * step (file:line in the reference, fbartolic/caustics):
 *   parameter map, x_cm shift ..... src/caustics/point_source.py:1798-1817
 *   _poly_coeffs_binary ........... src/caustics/point_source.py:17-86 (expanded monomials)
 *   _poly_coeffs_triple ........... src/caustics/point_source.py:89-1479 (the same polynomial
 *                                   built from the product form of the lens equation, as
 *                                   oracle/lens.py: poly_coeffs_product builds it)
 *   roots ......................... ora_ehrlich_aberth (ea_oracle.c: ehrlich_aberth.h:60-150,
 *                                   reference initial estimates), itmax 2500
 *   lens_eq, filter 1e-6 .......... src/caustics/point_source.py:1536-1556, 1652-1709
 *   lens_eq_det_jac, sum 1/|det| .. src/caustics/point_source.py:1558-1580, 1762-1830
 * Complex arithmetic is C99's (the oracle's NumPy restatement differs from it only in
 * the rounding of complex division), compiled without FMA contraction. */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;   /* the layout of ea_oracle.c's cplx */
int ora_ehrlich_aberth(int deg, int itmax, int compensated, int custom_init, int init_mode,
                       const cplx *poly, const cplx *roots_init, cplx *roots, int64_t *stats);

typedef double complex dc;

/* point_source.py:17-86 as oracle/lens.py restates it, highest-first */
static void coeffs_binary(dc w, double a, double e1, dc *p) {
    const dc wb = conj(w);
    /* integer powers as NumPy evaluates them: a**2 by a square, higher ones by pow */
    const double a2 = a * a, a3 = pow(a, 3.0), a4 = pow(a, 4.0), a5 = pow(a, 5.0), a6 = pow(a, 6.0);
    const dc wb2 = wb * wb;
    p[0] = -a2 + wb2;
    p[1] = a2 * w - 2 * a * e1 + a - w * wb2 + wb;
    p[2] = 2 * a4 - 2 * a2 * wb2 + 4 * a * wb * e1 - 2 * a * wb - 2 * w * wb;
    p[3] = (-2 * a4 * w + 4 * a3 * e1 - 2 * a3 + 2 * a2 * w * wb2 - 4 * a * w * wb * e1 +
            2 * a * w * wb + 2 * a * e1 - a - w);
    p[4] = (-a6 + a4 * wb2 - 4 * a3 * wb * e1 + 2 * a3 * wb + 2 * a2 * w * wb +
            4 * a2 * e1 * e1 - 4 * a2 * e1 + 2 * a2 - 4 * a * w * e1 + 2 * a * w);
    p[5] = (a6 * w - 2 * a5 * e1 + a5 - a4 * w * wb2 - a4 * wb + 4 * a3 * w * wb * e1 -
            2 * a3 * w * wb + 2 * a3 * e1 - a3 - 4 * a2 * w * e1 * e1 + 4 * a2 * w * e1 - a2 * w);
}

/* out = x * y, ascending coefficients */
static int pmul(const dc *x, int nx, const dc *y, int ny, dc *out) {
    for (int i = 0; i < nx + ny - 1; ++i) out[i] = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) out[i + j] += x[i] * y[j];
    return nx + ny - 1;
}

/* oracle/lens.py: poly_coeffs_product for m = 3 lenses, ascending (degree 10):
 * P(z) = (z - w) prod_j M_j - D sum_j e_j prod_{k!=j} M_k with D = prod_j (z - r_j),
 * N = sum_j e_j prod_{k!=j} (z - r_k), M_j = (conj(w) - conj(r_j)) D + N */
static void coeffs_triple(dc w, const dc *r, const double *e, dc *P) {
    dc lin[3][2], D[4], t2[3], tmp[4], Nn[3], M[3][4], M01[7], MM[10], zw[2], Pf[11], S[7],
        DS[10], Mk[7];
    for (int j = 0; j < 3; ++j) {
        lin[j][0] = -r[j];
        lin[j][1] = 1;
    }
    pmul(lin[0], 2, lin[1], 2, t2);
    pmul(t2, 3, lin[2], 2, D);
    /* sums of e_j prod_{k!=j}, each product associated as NumPy's (e_j * f_k1) * f_k2 */
    dc ef[4];
    for (int i = 0; i < 3; ++i) Nn[i] = 0;
    for (int j = 0; j < 3; ++j) {
        const int k1 = j == 0 ? 1 : 0, k2 = j == 2 ? 1 : 2;
        for (int i = 0; i < 2; ++i) ef[i] = e[j] * lin[k1][i];
        pmul(ef, 2, lin[k2], 2, tmp);
        for (int i = 0; i < 3; ++i) Nn[i] += tmp[i];
    }
    const dc wb = conj(w);
    for (int j = 0; j < 3; ++j) {
        const dc f = wb - conj(r[j]);
        for (int i = 0; i < 4; ++i) M[j][i] = f * D[i] + (i < 3 ? Nn[i] : 0);
    }
    pmul(M[0], 4, M[1], 4, M01);
    pmul(M01, 7, M[2], 4, MM);
    zw[0] = -w;
    zw[1] = 1;
    pmul(zw, 2, MM, 10, Pf);
    for (int i = 0; i < 7; ++i) S[i] = 0;
    for (int j = 0; j < 3; ++j) {
        const int k1 = j == 0 ? 1 : 0, k2 = j == 2 ? 1 : 2;
        for (int i = 0; i < 4; ++i) ef[i] = e[j] * M[k1][i];
        pmul(ef, 4, M[k2], 4, Mk);
        for (int i = 0; i < 7; ++i) S[i] += Mk[i];
    }
    pmul(D, 4, S, 7, DS);
    for (int i = 0; i < 11; ++i) P[i] = Pf[i] - (i < 10 ? DS[i] : 0);
}

/* one source position (user frame): the magnification; *ok = 0 when the solve did not
 * converge */
static double mag_one(int nlenses, dc w, const double *prm, int itmax, int *ok) {
    const double s = prm[0], q = prm[1], a = 0.5 * s, x_cm = a * (1 - q) / (1 + q);
    const dc wl = w + x_cm;
    dc r[3];
    double e[3];
    int deg, m;
    dc asc[11];
    if (nlenses == 2) {
        const double e1 = 1 / (1 + q);
        r[0] = a;
        r[1] = -a;
        e[0] = e1;
        e[1] = 1.0 - e1;
        m = 2;
        deg = 5;
        dc hi[6];
        coeffs_binary(wl, a, e1, hi);
        for (int i = 0; i <= deg; ++i) asc[i] = hi[deg - i];
    } else {
        const double q3 = prm[2], r3 = prm[3], psi = prm[4];
        const double e1 = q / (1 + q + q3), e2 = q * e1;
        r[0] = a;
        r[1] = -a;
        r[2] = r3 * cexp(I * psi);
        e[0] = e1;
        e[1] = e2;
        e[2] = 1.0 - e1 - e2;
        m = 3;
        deg = 10;
        coeffs_triple(wl, r, e, asc);
    }
    cplx poly[11], roots[10];
    memcpy(poly, asc, sizeof(cplx) * (deg + 1));
    *ok = ora_ehrlich_aberth(deg, itmax, 0, 0, 0, poly, NULL, roots, NULL);
    double mag = 0.0;
    for (int i = 0; i < deg; ++i) {
        const dc z = roots[i].re + I * roots[i].im, zb = conj(z);
        dc le = z, k = 0;
        for (int j = 0; j < m; ++j) {
            le -= e[j] / (zb - conj(r[j]));
            const dc d = zb - conj(r[j]);
            k += e[j] / (d * d);
        }
        if (cabs(le - wl) < 1e-6) {
            const double ak = cabs(k);
            mag += 1.0 / fabs(1.0 - ak * ak);
        }
    }
    return mag;
}

/* mag[i] for source positions w (N interleaved complex, the reference's frame) and
 * parameter rows params + i * pstride: (s, q) for two lenses, (s, q, q3, r3, psi) for
 * three.  nthreads <= 0: the OpenMP default.  Returns the number of unconverged solves. */
int64_t ora_mag_point_source(int nlenses, int64_t N, const double *w, const double *params,
                             int64_t pstride, double *mag, int itmax, int nthreads) {
    if (nlenses != 2 && nlenses != 3) return -1;
    int64_t bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
#endif
    for (int64_t i = 0; i < N; ++i) {
        int ok = 1;
        mag[i] = mag_one(nlenses, w[2 * i] + I * w[2 * i + 1], params + i * pstride, itmax, &ok);
        bad += !ok;
    }
    return bad;
}
