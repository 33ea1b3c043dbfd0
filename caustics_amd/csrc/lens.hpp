// Lens-plane arithmetic on the device: parameter map, lens polynomial, lens
// equation, Jacobian determinant and the implicit image derivative.
//
// Reference (fbartolic/caustics, src/caustics/point_source.py):
//   parameter map + x_cm shift ...... :1798-1817 (two lenses: a = s/2, e1 = 1/(1+q);
//                                      three: e1 = q/(1+q+q3), e2 = q e1, r3 e^{+i psi})
//   lens polynomial ................. :17-86 (two lenses), :89-1479 (three lenses)
//   lens_eq, lens_eq_det_jac ........ :1536-1580
//   image filter |lens_eq(z)-w|<1e-6  :1703-1707
//   magnification sum ............... :1828-1830
#pragma once
#include "cplx.hpp"

namespace caus {

template <int NL>
struct LensFrame {
    double e[NL];  // masses
    cd r[NL];      // positions
};

// Tangent of the lens frame along one parameter direction.
template <int NL>
struct LensTangent {
    double de[NL];
    cd dr[NL];
    double dxcm;
};

template <int NL>
constexpr int n_user_params() {
    return NL == 2 ? 2 : 5;
}

// User parameters (s, q) or (s, q, q3, r3, psi) -> lens frame and x_cm.
template <int NL>
__device__ __forceinline__ void lens_frame(const double* prm, LensFrame<NL>& lf, double& x_cm) {
    const double s = prm[0], q = prm[1];
    const double a = 0.5 * s;
    x_cm = a * (1.0 - q) / (1.0 + q);
    lf.r[0] = cd{a, 0.0};
    lf.r[1] = cd{-a, 0.0};
    if constexpr (NL == 2) {
        lf.e[0] = 1.0 / (1.0 + q);
        lf.e[1] = 1.0 - lf.e[0];
    } else {
        const double q3 = prm[2], R = prm[3], psi = prm[4];
        lf.e[0] = q / (1.0 + q + q3);
        lf.e[1] = q * lf.e[0];
        lf.e[2] = 1.0 - lf.e[0] - lf.e[1];
        double sp, cp;
        sincos(psi, &sp, &cp);
        lf.r[2] = cd{R * cp, R * sp};
    }
}

// d(lens frame)/d(user parameter k) for k < n_user_params<NL>().
template <int NL>
__device__ __forceinline__ void lens_frame_tangents(const double* prm,
                                                    LensTangent<NL> (&t)[n_user_params<NL>()]) {
    constexpr int NP = n_user_params<NL>();
    const double s = prm[0], q = prm[1];
    const double a = 0.5 * s;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            t[k].de[j] = 0.0;
            t[k].dr[j] = cd{0.0, 0.0};
        }
        t[k].dxcm = 0.0;
    }
    const double opq = 1.0 + q;
    // s
    t[0].dr[0] = cd{0.5, 0.0};
    t[0].dr[1] = cd{-0.5, 0.0};
    t[0].dxcm = 0.5 * (1.0 - q) / opq;
    // q
    t[1].dxcm = -2.0 * a / (opq * opq);
    if constexpr (NL == 2) {
        t[1].de[0] = -1.0 / (opq * opq);
        t[1].de[1] = -t[1].de[0];
    } else {
        const double q3 = prm[2], R = prm[3], psi = prm[4];
        const double den = 1.0 + q + q3;
        const double e1 = q / den;
        const double de1_dq = (1.0 + q3) / (den * den);
        const double de1_dq3 = -q / (den * den);
        t[1].de[0] = de1_dq;
        t[1].de[1] = e1 + q * de1_dq;
        t[1].de[2] = -t[1].de[0] - t[1].de[1];
        t[2].de[0] = de1_dq3;
        t[2].de[1] = q * de1_dq3;
        t[2].de[2] = -t[2].de[0] - t[2].de[1];
        double sp, cp;
        sincos(psi, &sp, &cp);
        t[3].dr[2] = cd{cp, sp};            // d r3 / d|r3|
        t[4].dr[2] = cd{-R * sp, R * cp};   // d r3 / d psi
    }
}

// ---- polynomials with compile-time sizes (ascending coefficients) ------------
template <int N>
struct Poly {
    cd c[N];
};
template <int A, int B>
__device__ __forceinline__ Poly<A + B - 1> pmul(const Poly<A>& x, const Poly<B>& y) {
    Poly<A + B - 1> o;
#pragma unroll
    for (int i = 0; i < A + B - 1; ++i) o.c[i] = cd{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
        for (int j = 0; j < B; ++j) o.c[i + j] = cfma(x.c[i], y.c[j], o.c[i + j]);
    return o;
}

// Lens polynomial in the lens frame, ascending, degree NL^2+1:
//   P(z) = (z - w) prod_j M_j - D sum_j e_j prod_{k!=j} M_k,
//   D = prod_j (z - r_j),  N = sum_j e_j prod_{k!=j} (z - r_k),  M_j = (conj(w) - conj(r_j)) D + N.
// Term for term the polynomial of point_source.py:17-86 / :89-1479.
template <int NL>
__device__ __forceinline__ void lens_poly(cd w, const LensFrame<NL>& lf, cd (&p)[NL * NL + 2]) {
    Poly<2> lin[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        lin[j].c[0] = -lf.r[j];
        lin[j].c[1] = cd{1.0, 0.0};
    }
    const cd wb = conj(w);
    if constexpr (NL == 2) {
        Poly<3> D = pmul(lin[0], lin[1]);
        Poly<2> Nn;
        Nn.c[0] = lf.e[0] * lin[1].c[0] + lf.e[1] * lin[0].c[0];
        Nn.c[1] = cd{lf.e[0] + lf.e[1], 0.0};
        Poly<3> M[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const cd f = wb - conj(lf.r[j]);
            M[j].c[0] = cfma(f, D.c[0], Nn.c[0]);
            M[j].c[1] = cfma(f, D.c[1], Nn.c[1]);
            M[j].c[2] = f * D.c[2];
        }
        Poly<5> MM = pmul(M[0], M[1]);
        Poly<2> zw;
        zw.c[0] = -w;
        zw.c[1] = cd{1.0, 0.0};
        Poly<6> P = pmul(zw, MM);
        Poly<3> S;
#pragma unroll
        for (int i = 0; i < 3; ++i) S.c[i] = lf.e[0] * M[1].c[i] + lf.e[1] * M[0].c[i];
        Poly<5> DS = pmul(D, S);
#pragma unroll
        for (int i = 0; i < 5; ++i) P.c[i] = P.c[i] - DS.c[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) p[i] = P.c[i];
    } else {
        Poly<3> D01 = pmul(lin[0], lin[1]);
        Poly<4> D = pmul(D01, lin[2]);
        Poly<3> D02 = pmul(lin[0], lin[2]);
        Poly<3> D12 = pmul(lin[1], lin[2]);
        Poly<3> Nn;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            Nn.c[i] = lf.e[0] * D12.c[i] + lf.e[1] * D02.c[i] + lf.e[2] * D01.c[i];
        Poly<4> M[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const cd f = wb - conj(lf.r[j]);
#pragma unroll
            for (int i = 0; i < 3; ++i) M[j].c[i] = cfma(f, D.c[i], Nn.c[i]);
            M[j].c[3] = f * D.c[3];
        }
        Poly<7> M01 = pmul(M[0], M[1]);
        Poly<7> M02 = pmul(M[0], M[2]);
        Poly<7> M12 = pmul(M[1], M[2]);
        Poly<10> MM = pmul(M01, M[2]);
        Poly<2> zw;
        zw.c[0] = -w;
        zw.c[1] = cd{1.0, 0.0};
        Poly<11> P = pmul(zw, MM);
        Poly<7> S;
#pragma unroll
        for (int i = 0; i < 7; ++i)
            S.c[i] = lf.e[0] * M12.c[i] + lf.e[1] * M02.c[i] + lf.e[2] * M01.c[i];
        Poly<10> DS = pmul(D, S);
#pragma unroll
        for (int i = 0; i < 10; ++i) P.c[i] = P.c[i] - DS.c[i];
#pragma unroll
        for (int i = 0; i < 11; ++i) p[i] = P.c[i];
    }
}

// Per image: G = sum_j e_j/(conj z - conj r_j) (so lens_eq(z) = z - G), K = sum_j
// e_j/(conj z - conj r_j)^2 (det J = 1 - |K|^2), K3 = sum_j e_j/(conj z - conj r_j)^3.
template <int NL>
struct ImageTerms {
    cd inv[NL];  // 1/(conj z - conj r_j)
    cd G, K, K3;
};

template <int NL>
__device__ __forceinline__ ImageTerms<NL> image_terms(cd z, const LensFrame<NL>& lf) {
    ImageTerms<NL> t;
    t.G = t.K = t.K3 = cd{0.0, 0.0};
    const cd zb = conj(z);
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = crcp<2>(zb - conj(lf.r[j]));
        const cd v2 = v * v;
        t.inv[j] = v;
        t.G = t.G + lf.e[j] * v;
        t.K = t.K + lf.e[j] * v2;
        t.K3 = t.K3 + lf.e[j] * (v2 * v);
    }
    return t;
}

// Change of one image's |det J|^-1 along a tangent direction with source-plane
// tangent dw (lens frame) and lens-frame tangent lt.  Implicit differentiation of
// w = z - G(conj z; theta):  dz + K conj(dz) = u,  u = dw + dG_theta
//   => dz = (u - K conj(u)) / det,   dK = -2 K3 conj(dz) + dK_theta,
//   d det = -2 Re(conj(K) dK),       d(1/|det|) = -sign(det) d det / det^2.
template <int NL>
__device__ __forceinline__ cd image_dz(const ImageTerms<NL>& it, const LensFrame<NL>& lf,
                                       double det, cd dw, const LensTangent<NL>& lt) {
    cd dG = cd{0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = it.inv[j];
        dG = dG + lt.de[j] * v;
        dG = cfma(lf.e[j] * (v * v), conj(lt.dr[j]), dG);
    }
    const cd u = dw + dG;
    return rcp(det) * (u - it.K * conj(u));
}

template <int NL>
__device__ __forceinline__ double dmu_image(const ImageTerms<NL>& it, const LensFrame<NL>& lf,
                                            double det, cd dw, const LensTangent<NL>& lt) {
    cd dG = cd{0.0, 0.0}, dK = cd{0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = it.inv[j];
        const cd v2 = v * v;
        // d/d(mass)
        dG = dG + lt.de[j] * v;
        dK = dK + lt.de[j] * v2;
        // d/d(conj r_j): G -> e v^2, K -> 2 e v^3
        const cd drb = conj(lt.dr[j]);
        dG = cfma(lf.e[j] * v2, drb, dG);
        dK = cfma((2.0 * lf.e[j]) * (v2 * v), drb, dK);
    }
    const cd u = dw + dG;
    const double idet = rcp(det);
    const cd dz = idet * (u - it.K * conj(u));
    dK = dK - 2.0 * (it.K3 * conj(dz));
    const double ddet = -2.0 * (it.K.re * dK.re + it.K.im * dK.im);
    return -copysign(1.0, det) * ddet * idet * idet;
}

}  // namespace caus
