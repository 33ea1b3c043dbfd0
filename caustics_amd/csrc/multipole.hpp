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

// a_pq = mu0 (conj(Q) + conj(W2) Q)
__device__ __forceinline__ cd akl(double mu0, cd W2c, cd Q) { return mu0 * (conj(Q) + W2c * Q); }

__device__ __forceinline__ Multipole hexadecapole_cassan(cd W2, cd W3, cd W4, cd W5, cd W6,
                                                        double rho, double u1) {
    const double Gamma = 2.0 * u1 / (3.0 - u1);
    const cd W2c = conj(W2);
    const double mu0 = 1.0 / (1.0 - norm2(W2));
    const cd one = cd{1.0, 0.0};
    const cd a10 = mu0 * (one + W2c);
    const cd a01 = mu0 * (cd{0.0, 1.0} * (one - W2c));
    // order 2
    const cd a20 = akl(mu0, W2c, W3 * a10 * a10);
    const cd a11 = akl(mu0, W2c, W3 * a10 * a01);
    const cd a02 = akl(mu0, W2c, W3 * a01 * a01);
    const cd a10_2 = a10 * a10, a01_2 = a01 * a01;
    const cd a10_3 = a10_2 * a10, a01_3 = a01_2 * a01;
    // order 3
    const cd a30 = akl(mu0, W2c, W3 * (3.0 * a20 * a10) + W4 * a10_3);
    const cd a21 = akl(mu0, W2c, W3 * (2.0 * a11 * a10 + a01 * a20) + W4 * a01 * a10_2);
    const cd a12 = akl(mu0, W2c, W3 * (2.0 * a11 * a01 + a10 * a02) + W4 * a10 * a01_2);
    const cd a03 = akl(mu0, W2c, W3 * (3.0 * a02 * a01) + W4 * a01_3);
    // order 4
    const cd a40 = akl(mu0, W2c,
                       W3 * (4.0 * a30 * a10 + 3.0 * a20 * a20) + W4 * (6.0 * a20 * a10_2) +
                           W5 * (a10_3 * a10));
    const cd a31 = akl(mu0, W2c,
                       W3 * (3.0 * a21 * a10 + 3.0 * a11 * a20 + a01 * a30) +
                           W4 * (3.0 * (a11 * a10_2 + a01 * a20 * a10)) + W5 * (a01 * a10_3));
    const cd a22 = akl(mu0, W2c,
                       W3 * (2.0 * a12 * a10 + 2.0 * a11 * a11 + a02 * a20 + 2.0 * a21 * a01) +
                           W4 * (a02 * a10_2 + 4.0 * a11 * a01 * a10 + a01_2 * a20) +
                           W5 * (a01_2 * a10_2));
    const cd a13 = akl(mu0, W2c,
                       W3 * (3.0 * a12 * a01 + 3.0 * a11 * a02 + a10 * a03) +
                           W4 * (3.0 * (a11 * a01_2 + a10 * a02 * a01)) + W5 * (a10 * a01_3));
    const cd a04 = akl(mu0, W2c,
                       W3 * (4.0 * a03 * a01 + 3.0 * a02 * a02) + W4 * (6.0 * a02 * a01_2) +
                           W5 * (a01_3 * a01));
    // order 5
    const cd a50 = akl(mu0, W2c,
                       W3 * (5.0 * a40 * a10 + 10.0 * a20 * a30) +
                           W4 * (10.0 * a30 * a10_2 + 15.0 * a20 * a20 * a10) +
                           W5 * (10.0 * a20 * a10_3) + W6 * (a10_3 * a10_2));
    const cd a41 = akl(mu0, W2c,
                       W3 * (4.0 * a31 * a10 + 4.0 * a11 * a30 + 6.0 * a21 * a20 + a01 * a40) +
                           W4 * (6.0 * a21 * a10_2 + 12.0 * a11 * a20 * a10 +
                                 3.0 * a01 * a20 * a20 + 4.0 * a01 * a10 * a30) +
                           W5 * (4.0 * a11 * a10_3 + 6.0 * a01 * a10_2 * a20) +
                           W6 * (a01 * a10_3 * a10));
    const cd a32 = akl(mu0, W2c,
                       W3 * (3.0 * a22 * a10 + 6.0 * a11 * a21 + 3.0 * a12 * a20 + a02 * a30 +
                             2.0 * a31 * a01) +
                           W4 * (3.0 * a12 * a10_2 + 6.0 * a11 * a11 * a10 +
                                 3.0 * a02 * a20 * a10 + 6.0 * a01 * a11 * a20 +
                                 6.0 * a01 * a21 * a10 + a01_2 * a30) +
                           W5 * (a02 * a10_3 + 6.0 * a11 * a01 * a10_2 + 3.0 * a01_2 * a10 * a20) +
                           W6 * (a01_2 * a10_3));
    const cd a23 = akl(mu0, W2c,
                       W3 * (3.0 * a22 * a01 + 6.0 * a11 * a12 + 3.0 * a21 * a02 + a20 * a03 +
                             2.0 * a13 * a10) +
                           W4 * (3.0 * a21 * a01_2 + 6.0 * a11 * a11 * a01 +
                                 3.0 * a20 * a02 * a01 + 6.0 * a10 * a11 * a02 +
                                 6.0 * a10 * a12 * a01 + a10_2 * a03) +
                           W5 * (a20 * a01_3 + 6.0 * a11 * a10 * a01_2 + 3.0 * a10_2 * a01 * a02) +
                           W6 * (a10_2 * a01_3));
    const cd a14 = akl(mu0, W2c,
                       W3 * (4.0 * a13 * a01 + 4.0 * a11 * a03 + 6.0 * a12 * a02 + a10 * a04) +
                           W4 * (6.0 * a12 * a01_2 + 12.0 * a11 * a02 * a01 +
                                 3.0 * a10 * a02 * a02 + 4.0 * a10 * a01 * a03) +
                           W5 * (4.0 * a11 * a01_3 + 6.0 * a10 * a01_2 * a02) +
                           W6 * (a10 * a01_3 * a01));
    const cd a05 = akl(mu0, W2c,
                       W3 * (5.0 * a04 * a01 + 10.0 * a02 * a03) +
                           W4 * (10.0 * a03 * a01_2 + 15.0 * a02 * a02 * a01) +
                           W5 * (10.0 * a02 * a01_3) + W6 * (a01_3 * a01_2));
    const cd m2 = a01 * conj(a12 + a30) + conj(a10) * (a03 + a21) + 2.0 * (a02 * conj(a11)) +
                  2.0 * (a11 * conj(a20));
    const cd m4 = (a05 + 2.0 * a23 + a41) * conj(a10) + 4.0 * (a04 * conj(a11)) +
                  4.0 * ((a13 + a31) * conj(a20)) + 6.0 * (a12 * conj(a21)) +
                  6.0 * (a21 * conj(a30)) + a03 * (6.0 * conj(a12) + 2.0 * conj(a30)) +
                  4.0 * (a02 * conj(a13 + a31)) + 4.0 * (a11 * conj(a40)) +
                  a01 * conj(a14 + 2.0 * a32 + a50);
    const double mu2 = 0.25 * m2.im, mu4 = 0.125 * m4.im;
    Multipole r;
    r.mu0 = mu0;
    r.dq = 0.5 * mu2 * (1.0 - 0.2 * Gamma) * rho * rho;
    r.dh = mu4 * (1.0 - 11.0 / 35.0 * Gamma) * (rho * rho) * (rho * rho) / 24.0;
    return r;
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

}  // namespace caus
