// Hexadecapole approximation (Cassan 2017) and the contour gate, per image on device.
//
// Reference (fbartolic/caustics):
//   _mag_hexadecapole_cassan ... src/caustics/multipole.py:8-205 (the explicit order-2..5
//                                recurrences for a_pq; the same published expressions)
//   _mag_hexadecapole .......... src/caustics/multipole.py:208-244
//   _caustics_proximity_test ... src/caustics/lightcurve.py:23-75
//   _planetary_caustic_test .... src/caustics/lightcurve.py:78-85
#pragma once
#include "cplx.hpp"

namespace caus {

__device__ __forceinline__ cd cpow_i(cd x, int k) {
    cd r = x;
    for (int i = 1; i < k; ++i) r = r * x;
    return r;
}

struct Multipole {
    double mu0, dq, dh;
};

// First-order dual numbers (value and one tangent) for the forward-mode derivative
// of the multipole expansion: the same source text below evaluates either plain
// values (cd, double) or values with tangents (dcd, ddl).
struct ddl {
    double v, d;
};
struct dcd {
    cd v, d;
};
__device__ __forceinline__ dcd operator+(const dcd& a, const dcd& b) { return dcd{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dcd operator-(const dcd& a, const dcd& b) { return dcd{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dcd operator*(const dcd& a, const dcd& b) {
    return dcd{a.v * b.v, cfma(a.d, b.v, a.v * b.d)};
}
__device__ __forceinline__ dcd operator*(double s, const dcd& a) { return dcd{s * a.v, s * a.d}; }
__device__ __forceinline__ dcd operator*(const ddl& s, const dcd& a) {
    return dcd{s.v * a.v, cfma(cd{s.d, 0.0}, a.v, s.v * a.d)};
}
__device__ __forceinline__ dcd conj(const dcd& a) { return dcd{conj(a.v), conj(a.d)}; }
__device__ __forceinline__ dcd mul_i(const dcd& a) {
    return dcd{cd{-a.v.im, a.v.re}, cd{-a.d.im, a.d.re}};
}
__device__ __forceinline__ cd mul_i(cd a) { return cd{-a.im, a.re}; }
__device__ __forceinline__ ddl imag_part(const dcd& a) { return ddl{a.v.im, a.d.im}; }
__device__ __forceinline__ double imag_part(cd a) { return a.im; }
// 1 / (1 - |W|^2)
__device__ __forceinline__ double inv_one_minus_norm2(cd W) { return 1.0 / (1.0 - norm2(W)); }
__device__ __forceinline__ ddl inv_one_minus_norm2(const dcd& W) {
    const double m = 1.0 / (1.0 - norm2(W.v));
    return ddl{m, 2.0 * m * m * (W.v.re * W.d.re + W.v.im * W.d.im)};
}
template <class C>
__device__ __forceinline__ C cconst(double re);
template <>
__device__ __forceinline__ cd cconst<cd>(double re) { return cd{re, 0.0}; }
template <>
__device__ __forceinline__ dcd cconst<dcd>(double re) { return dcd{cd{re, 0.0}, cd{0.0, 0.0}}; }

// a_pq = mu0 (conj(Q) + conj(W2) Q)
template <class C, class R>
__device__ __forceinline__ C akl(const R& mu0, const C& W2c, const C& Q) {
    return mu0 * (conj(Q) + W2c * Q);
}

// mu0 and the imaginary parts mu2, mu4 of the Cassan (2017) moment sums
// (multipole.py:25-199) from W2..W6.
template <class C, class R>
__device__ __forceinline__ void cassan_moments(const C& W2, const C& W3, const C& W4,
                                               const C& W5, const C& W6, R& mu0, R& mu2,
                                               R& mu4) {
    const C W2c = conj(W2);
    mu0 = inv_one_minus_norm2(W2);
    const C one = cconst<C>(1.0);
    const C a10 = mu0 * (one + W2c);
    const C a01 = mu0 * mul_i(one - W2c);
    // order 2
    const C a20 = akl(mu0, W2c, W3 * a10 * a10);
    const C a11 = akl(mu0, W2c, W3 * a10 * a01);
    const C a02 = akl(mu0, W2c, W3 * a01 * a01);
    const C a10_2 = a10 * a10, a01_2 = a01 * a01;
    const C a10_3 = a10_2 * a10, a01_3 = a01_2 * a01;
    // order 3
    const C a30 = akl(mu0, W2c, W3 * (3.0 * a20 * a10) + W4 * a10_3);
    const C a21 = akl(mu0, W2c, W3 * (2.0 * a11 * a10 + a01 * a20) + W4 * a01 * a10_2);
    const C a12 = akl(mu0, W2c, W3 * (2.0 * a11 * a01 + a10 * a02) + W4 * a10 * a01_2);
    const C a03 = akl(mu0, W2c, W3 * (3.0 * a02 * a01) + W4 * a01_3);
    // order 4
    const C a40 = akl(mu0, W2c,
                      W3 * (4.0 * a30 * a10 + 3.0 * a20 * a20) + W4 * (6.0 * a20 * a10_2) +
                          W5 * (a10_3 * a10));
    const C a31 = akl(mu0, W2c,
                      W3 * (3.0 * a21 * a10 + 3.0 * a11 * a20 + a01 * a30) +
                          W4 * (3.0 * (a11 * a10_2 + a01 * a20 * a10)) + W5 * (a01 * a10_3));
    const C a22 = akl(mu0, W2c,
                      W3 * (2.0 * a12 * a10 + 2.0 * a11 * a11 + a02 * a20 + 2.0 * a21 * a01) +
                          W4 * (a02 * a10_2 + 4.0 * a11 * a01 * a10 + a01_2 * a20) +
                          W5 * (a01_2 * a10_2));
    const C a13 = akl(mu0, W2c,
                      W3 * (3.0 * a12 * a01 + 3.0 * a11 * a02 + a10 * a03) +
                          W4 * (3.0 * (a11 * a01_2 + a10 * a02 * a01)) + W5 * (a10 * a01_3));
    const C a04 = akl(mu0, W2c,
                      W3 * (4.0 * a03 * a01 + 3.0 * a02 * a02) + W4 * (6.0 * a02 * a01_2) +
                          W5 * (a01_3 * a01));
    // order 5
    const C a50 = akl(mu0, W2c,
                      W3 * (5.0 * a40 * a10 + 10.0 * a20 * a30) +
                          W4 * (10.0 * a30 * a10_2 + 15.0 * a20 * a20 * a10) +
                          W5 * (10.0 * a20 * a10_3) + W6 * (a10_3 * a10_2));
    const C a41 = akl(mu0, W2c,
                      W3 * (4.0 * a31 * a10 + 4.0 * a11 * a30 + 6.0 * a21 * a20 + a01 * a40) +
                          W4 * (6.0 * a21 * a10_2 + 12.0 * a11 * a20 * a10 +
                                3.0 * a01 * a20 * a20 + 4.0 * a01 * a10 * a30) +
                          W5 * (4.0 * a11 * a10_3 + 6.0 * a01 * a10_2 * a20) +
                          W6 * (a01 * a10_3 * a10));
    const C a32 = akl(mu0, W2c,
                      W3 * (3.0 * a22 * a10 + 6.0 * a11 * a21 + 3.0 * a12 * a20 + a02 * a30 +
                            2.0 * a31 * a01) +
                          W4 * (3.0 * a12 * a10_2 + 6.0 * a11 * a11 * a10 +
                                3.0 * a02 * a20 * a10 + 6.0 * a01 * a11 * a20 +
                                6.0 * a01 * a21 * a10 + a01_2 * a30) +
                          W5 * (a02 * a10_3 + 6.0 * a11 * a01 * a10_2 + 3.0 * a01_2 * a10 * a20) +
                          W6 * (a01_2 * a10_3));
    const C a23 = akl(mu0, W2c,
                      W3 * (3.0 * a22 * a01 + 6.0 * a11 * a12 + 3.0 * a21 * a02 + a20 * a03 +
                            2.0 * a13 * a10) +
                          W4 * (3.0 * a21 * a01_2 + 6.0 * a11 * a11 * a01 +
                                3.0 * a20 * a02 * a01 + 6.0 * a10 * a11 * a02 +
                                6.0 * a10 * a12 * a01 + a10_2 * a03) +
                          W5 * (a20 * a01_3 + 6.0 * a11 * a10 * a01_2 + 3.0 * a10_2 * a01 * a02) +
                          W6 * (a10_2 * a01_3));
    const C a14 = akl(mu0, W2c,
                      W3 * (4.0 * a13 * a01 + 4.0 * a11 * a03 + 6.0 * a12 * a02 + a10 * a04) +
                          W4 * (6.0 * a12 * a01_2 + 12.0 * a11 * a02 * a01 +
                                3.0 * a10 * a02 * a02 + 4.0 * a10 * a01 * a03) +
                          W5 * (4.0 * a11 * a01_3 + 6.0 * a10 * a01_2 * a02) +
                          W6 * (a10 * a01_3 * a01));
    const C a05 = akl(mu0, W2c,
                      W3 * (5.0 * a04 * a01 + 10.0 * a02 * a03) +
                          W4 * (10.0 * a03 * a01_2 + 15.0 * a02 * a02 * a01) +
                          W5 * (10.0 * a02 * a01_3) + W6 * (a01_3 * a01_2));
    const C m2 = a01 * conj(a12 + a30) + conj(a10) * (a03 + a21) + 2.0 * (a02 * conj(a11)) +
                 2.0 * (a11 * conj(a20));
    const C m4 = (a05 + 2.0 * a23 + a41) * conj(a10) + 4.0 * (a04 * conj(a11)) +
                 4.0 * ((a13 + a31) * conj(a20)) + 6.0 * (a12 * conj(a21)) +
                 6.0 * (a21 * conj(a30)) + a03 * (6.0 * conj(a12) + 2.0 * conj(a30)) +
                 4.0 * (a02 * conj(a13 + a31)) + 4.0 * (a11 * conj(a40)) +
                 a01 * conj(a14 + 2.0 * a32 + a50);
    mu2 = imag_part(m2);
    mu4 = imag_part(m4);
    // mu2 = 0.25 Im(m2), mu4 = 0.125 Im(m4) (multipole.py:175-199): scaled by the caller
}

__device__ __forceinline__ Multipole hexadecapole_cassan(cd W2, cd W3, cd W4, cd W5, cd W6,
                                                        double rho, double u1) {
    const double Gamma = 2.0 * u1 / (3.0 - u1);
    double mu0, m2, m4;
    cassan_moments<cd, double>(W2, W3, W4, W5, W6, mu0, m2, m4);
    const double mu2 = 0.25 * m2, mu4 = 0.125 * m4;
    Multipole r;
    r.mu0 = mu0;
    r.dq = 0.5 * mu2 * (1.0 - 0.2 * Gamma) * rho * rho;
    r.dh = mu4 * (1.0 - 11.0 / 35.0 * Gamma) * (rho * rho) * (rho * rho) / 24.0;
    return r;
}

// Tangent of mu0 + dq + dh along a tangent dW of (W2..W6) at fixed rho, u1.
__device__ __forceinline__ double hexadecapole_tangent(const cd (&W)[5], const cd (&dW)[5],
                                                       double rho, double u1) {
    const double Gamma = 2.0 * u1 / (3.0 - u1);
    dcd Wd[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Wd[k] = dcd{W[k], dW[k]};
    ddl mu0, m2, m4;
    cassan_moments<dcd, ddl>(Wd[0], Wd[1], Wd[2], Wd[3], Wd[4], mu0, m2, m4);
    const double r2 = rho * rho;
    return mu0.d + 0.125 * m2.d * (1.0 - 0.2 * Gamma) * r2 +
           0.125 * m4.d * (1.0 - 11.0 / 35.0 * Gamma) * r2 * r2 / 24.0;
}

// W_k = (-1)^(k-1) (k-1)! sum_j e_j/(z - r_j)^k, k = 2..6 (multipole.py:212-231)
template <int NL>
__device__ __forceinline__ void w_terms(cd z, const double (&e)[NL], const cd (&r)[NL], cd (&W)[5]) {
    cd s[5] = {cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = crcp<2>(z - r[j]);
        cd pw = v * v;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            s[k] = s[k] + e[j] * pw;
            pw = pw * v;
        }
    }
    // (-1)^(k-1) (k-1)! for k = 2..6
    W[0] = -1.0 * s[0];
    W[1] = 2.0 * s[1];
    W[2] = -6.0 * s[2];
    W[3] = 24.0 * s[3];
    W[4] = -120.0 * s[4];
}

// Tangent of W_2..W_6 at an image z along (dz, de_j, dr_j):
//   dW_k = W_{k+1} dz + (-1)^(k-1) (k-1)! sum_j [de_j v_j^k + k e_j v_j^(k+1) dr_j],
//   v_j = 1/(z - r_j)  (since d/dz W_k = W_{k+1}).
template <int NL>
__device__ __forceinline__ void w_terms_tangent(cd z, const double (&e)[NL], const cd (&r)[NL],
                                                cd dz, const double (&de)[NL],
                                                const cd (&dr)[NL], cd (&dW)[5]) {
    const double c[6] = {-1.0, 2.0, -6.0, 24.0, -120.0, 720.0};  // (-1)^(k-1)(k-1)!, k=2..7
    cd sv[6], sr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) sv[k] = sr[k] = cd{0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = crcp<2>(z - r[j]);
        cd pw = v * v;  // v^2
#pragma unroll
        for (int k = 0; k < 6; ++k) {   // power k + 2
            sv[k] = sv[k] + e[j] * pw;  // sum e v^(k+2)
            if (k < 5) {
                // (k+2) e v^(k+3) dr  and  de v^(k+2)
                sr[k] = sr[k] + de[j] * pw + ((double)(k + 2) * e[j]) * (pw * v * dr[j]);
            }
            pw = pw * v;
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) dW[k] = c[k] * sr[k] + (c[k + 1] * sv[k + 1]) * dz;
}

}  // namespace caus
