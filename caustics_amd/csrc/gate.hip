// Hexadecapole magnification and the "needs a contour" gate of the light-curve
// function mag() (reference: src/caustics/lightcurve.py:168-227), one source point
// per lane: images, Cassan multipoles, caustic-proximity / ghost-image tests and the
// planetary-caustic test.
#include "caustics_hip.h"
#include "multipole.hpp"
#include "ps_impl.hpp"

using namespace caus;

namespace {

constexpr int kGateBlock = 256;

template <int NL>
__device__ __forceinline__ void lens_arrays(const LensFrame<NL>& lf, double (&e)[NL], cd (&r)[NL]) {
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        e[j] = lf.e[j];
        r[j] = lf.r[j];
    }
}

// f, f', f'' of the binary lens map used by the proximity test (lightcurve.py:27-31)
struct BinF {
    double a, e1;
    __device__ cd f(cd x) const {
        return -1.0 * (e1 * crcp<2>(x - cd{a, 0}) + (1.0 - e1) * crcp<2>(x + cd{a, 0}));
    }
    __device__ cd fp(cd x) const {
        const cd u = crcp<2>(x - cd{a, 0}), v = crcp<2>(x + cd{a, 0});
        return e1 * (u * u) + (1.0 - e1) * (v * v);
    }
    __device__ cd fpp(cd x) const {
        const cd u = crcp<2>(cd{a, 0} - x), v = crcp<2>(cd{a, 0} + x);
        return 2.0 * (e1 * (u * u * u) - (1.0 - e1) * (v * v * v));
    }
};

template <int NL>
__global__ __launch_bounds__(kGateBlock) void k_hexadecapole_gate(
    const cd* __restrict__ w, int64_t P, const double* __restrict__ params, int64_t pstride,
    const double* __restrict__ rho, int64_t rstride, double u1, double* __restrict__ mu_hex,
    double* __restrict__ dmu_out, uint8_t* __restrict__ use_hex, uint8_t* __restrict__ status,
    int itmax) {
    const int64_t i = (int64_t)blockIdx.x * kGateBlock + threadIdx.x;
    if (i >= P) return;
    const double* prm = params + i * pstride;
    const double rh = rho[i * rstride];
    LensFrame<NL> lf;
    double x_cm = 0.0;
    if constexpr (NL == 1) {
        lf.e[0] = 1.0;
        lf.r[0] = cd{0.0, 0.0};
    } else {
        lens_frame<NL>(prm, lf, x_cm);
    }
    const cd wl = w[i] + cd{x_cm, 0.0};
    constexpr int DEG = NL == 1 ? 2 : NL * NL + 1;
    cd z[DEG];
    bool real_img[DEG];
    bool ok = true;
    if constexpr (NL == 1) {
        const double sq = sqrt(1.0 + 4.0 * rcp(norm2(wl)));
        z[0] = (0.5 * (1.0 + sq)) * wl;
        z[1] = (0.5 * (1.0 - sq)) * wl;
        real_img[0] = real_img[1] = true;
    } else {
        ok = solve_lens<NL, MODE_FP64>(wl, lf, z, itmax);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            ImageTerms<NL> it = image_terms<NL>(z[j], lf);
            real_img[j] = norm2(z[j] - it.G - wl) < 1e-12;
        }
    }
    double e[NL];
    cd r[NL];
    lens_arrays<NL>(lf, e, r);
    double mu = 0.0, dq_abs = 0.0, dh_abs = 0.0;
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
        if (real_img[j]) {
            cd W[5];
            w_terms<NL>(z[j], e, r, W);
            const Multipole m = hexadecapole_cassan(W[0], W[1], W[2], W[3], W[4], rh, u1);
            mu += fabs(m.mu0 + m.dq + m.dh);
            dq_abs += fabs(m.dq);
            dh_abs += fabs(m.dh);
        }
    }
    const double dmu = dq_abs + dh_abs;
    bool use;
    if constexpr (NL == 1) {
        use = sqrt(norm2(wl)) > 2.0 * rh;
    } else if constexpr (NL == 3) {
        use = false;  // lightcurve.py:186-187, 226-227: full contour for three lenses
    } else {
        // _caustics_proximity_test with c_m=1e-2, gamma=0.02, c_f=4, rho_min=1e-3
        const double c_m = 1e-2, gamma = 0.02, c_f = 4.0, rho_min = 1e-3;
        const BinF F{lf.r[0].re, lf.e[0]};
        double mu_cusp = 0.0, ghost = 0.0;
        int nfalse = 0;
        const double rr = (rh + rho_min) * (rh + rho_min);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            const cd zb = conj(z[j]);
            const cd zh = conj(wl) - F.f(z[j]);
            const cd fp_z = F.fp(z[j]), fpp_z = F.fpp(z[j]), fp_zb = F.fp(zb), fp_zh = F.fp(zh);
            const cd fpp_zb = F.fpp(zb);
            const double J = 1.0 - sqrt(norm2(fp_z * fp_zb));
            if (real_img[j]) {
                const cd t = 3.0 * (fp_zb * fp_zb * fp_zb) * (fpp_z * fpp_z);
                const double J2 = J * J;
                mu_cusp += fabs(6.0 * t.im / (J2 * J2 * J) * rr);
            } else {
                ++nfalse;
                const double Jh = 1.0 - sqrt(norm2(fp_z * fp_zh));
                const cd den = Jh * (fpp_zb * fp_z) - Jh * (fpp_z * fp_zb * fp_zh);
                ghost += sqrt(norm2(cdiv<2>(cd{J * Jh * Jh, 0.0}, den)));
            }
        }
        const bool t_mc = gamma * mu_cusp + dmu < c_m;
        const bool t_fi = (nfalse == 0) ? true : (0.5 * ghost > c_f * (rh + rho_min));
        use = t_mc && t_fi;
        const double q = prm[1];
        if (q < 0.01) {
            // _planetary_caustic_test, c_p = 2; its q is e1/(1-e1)
            const double a = lf.r[0].re, e1 = lf.e[0];
            const double s = 2.0 * a, qq = e1 / (1.0 - e1);
            const double wpc = -1.0 / s, dpc = 3.0 * sqrt(qq) / s;
            const double dx = wpc - wl.re, dy = -wl.im;
            use = use && (dx * dx + dy * dy > 2.0 * (rh * rh + dpc * dpc));
        }
    }
    mu_hex[i] = mu;
    if (dmu_out) dmu_out[i] = dmu;
    use_hex[i] = use ? 1 : 0;
    if (status) status[i] = ok ? 1 : 0;
}

// Reverse mode of the hexadecapole magnification sum_images |mu0 + dq + dh|
// (multipole.py:208-244) with respect to (Re w, Im w, rho, u1, lens parameters), the
// derivative JAX's autodiff gives the gated points of mag() (lightcurve.py:215-254).
// Per real image: image tangents dz from the implicit lens equation, W_k tangents
// dW_k = W_{k+1} dz + d_theta W_k, then one forward-mode (dual-number) pass of the
// Cassan recurrences per direction; rho and u1 enter dq, dh in closed form.
// g_out row: [Re w, Im w, rho, u1, lens params...] times g_mu.
template <int NL>
__global__ __launch_bounds__(kGateBlock) void k_hexadecapole_vjp(
    const cd* __restrict__ w, int64_t P, const double* __restrict__ params, int64_t pstride,
    const double* __restrict__ rho, int64_t rstride, double u1, const double* __restrict__ g_mu,
    double* __restrict__ mu_hex, double* __restrict__ g_out, uint8_t* __restrict__ status,
    int itmax) {
    constexpr int NP = NL == 1 ? 0 : n_user_params<NL>();
    constexpr int NG = 4 + NP;
    constexpr int ND = 2 + NP;
    const int64_t i = (int64_t)blockIdx.x * kGateBlock + threadIdx.x;
    if (i >= P) return;
    const double* prm = params + i * pstride;
    const double rh = rho[i * rstride];
    LensFrame<NL> lf;
    double x_cm = 0.0;
    if constexpr (NL == 1) {
        lf.e[0] = 1.0;
        lf.r[0] = cd{0.0, 0.0};
    } else {
        lens_frame<NL>(prm, lf, x_cm);
    }
    const cd wl = w[i] + cd{x_cm, 0.0};
    // tangent directions: Re w, Im w, then each lens parameter
    cd dw[ND];
    LensTangent<NL> lt[ND];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            lt[d].de[j] = 0.0;
            lt[d].dr[j] = cd{0.0, 0.0};
        }
        lt[d].dxcm = 0.0;
    }
    dw[0] = cd{1.0, 0.0};
    dw[1] = cd{0.0, 1.0};
    if constexpr (NP > 0) {
        LensTangent<NL> up[NP];
        lens_frame_tangents<NL>(prm, up);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            lt[2 + k] = up[k];
            dw[2 + k] = cd{up[k].dxcm, 0.0};
        }
    }
    constexpr int DEG = NL == 1 ? 2 : NL * NL + 1;
    cd z[DEG];
    bool real_img[DEG];
    bool ok = true;
    if constexpr (NL == 1) {
        const double sq = sqrt(1.0 + 4.0 * rcp(norm2(wl)));
        z[0] = (0.5 * (1.0 + sq)) * wl;
        z[1] = (0.5 * (1.0 - sq)) * wl;
        real_img[0] = real_img[1] = true;
    } else {
        ok = solve_lens<NL, MODE_FP64>(wl, lf, z, itmax);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            ImageTerms<NL> it = image_terms<NL>(z[j], lf);
            real_img[j] = norm2(z[j] - it.G - wl) < 1e-12;
        }
    }
    double e[NL];
    cd r[NL];
    lens_arrays<NL>(lf, e, r);
    const double Gamma = 2.0 * u1 / (3.0 - u1);
    const double dGamma = 6.0 / ((3.0 - u1) * (3.0 - u1));
    const double r2 = rh * rh;
    double mu = 0.0, grho = 0.0, gu1 = 0.0;
    double gd[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) gd[d] = 0.0;
    for (int j = 0; j < DEG; ++j) {
        if (!real_img[j]) continue;
        cd W[5];
        w_terms<NL>(z[j], e, r, W);
        double mu0, m2, m4;
        cassan_moments<cd, double>(W[0], W[1], W[2], W[3], W[4], mu0, m2, m4);
        const double fq = 0.125 * m2, fh = 0.125 * m4 / 24.0;  // dq = fq (1-G/5) r^2, ...
        const double cq = 1.0 - 0.2 * Gamma, ch = 1.0 - 11.0 / 35.0 * Gamma;
        const double val = mu0 + fq * cq * r2 + fh * ch * r2 * r2;
        const double sg = copysign(1.0, val);
        mu += fabs(val);
        grho += sg * (2.0 * fq * cq * rh + 4.0 * fh * ch * r2 * rh);
        gu1 += sg * (-(0.2 * fq * r2 + 11.0 / 35.0 * fh * r2 * r2) * dGamma);
        const ImageTerms<NL> it = image_terms<NL>(z[j], lf);
        const double det = 1.0 - norm2(it.K);
#pragma unroll 1
        for (int d = 0; d < ND; ++d) {
            const cd dz = image_dz<NL>(it, lf, det, dw[d], lt[d]);
            cd dW[5];
            w_terms_tangent<NL>(z[j], e, r, dz, lt[d].de, lt[d].dr, dW);
            gd[d] += sg * hexadecapole_tangent(W, dW, rh, u1);
        }
    }
    mu_hex[i] = mu;
    if (status) status[i] = ok ? 1 : 0;
    const double g = g_mu[i];
    double* go = g_out + i * NG;
    go[0] = g * gd[0];
    go[1] = g * gd[1];
    go[2] = g * grho;
    go[3] = g * gu1;
#pragma unroll
    for (int k = 0; k < NP; ++k) go[4 + k] = g * gd[2 + k];
}

}  // namespace

extern "C" int caus_hexadecapole_gate(void* stream, int32_t nlenses, const double* w, int64_t P,
                                      const double* params, int64_t params_stride,
                                      const double* rho, int64_t rho_stride, double u1,
                                      double* mu_hex, double* delta_mu, uint8_t* use_hex,
                                      uint8_t* status, int32_t itmax) {
    if (P < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || params_stride < 0 || rho_stride < 0)
        return CAUS_ERR_ARG;
    if (P == 0) return CAUS_OK;
    if (!w || !rho || !mu_hex || !use_hex || (nlenses > 1 && !params)) return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((P + kGateBlock - 1) / kGateBlock)), block(kGateBlock);
    const cd* wc = reinterpret_cast<const cd*>(w);
    switch (nlenses) {
        case 1:
            hipLaunchKernelGGL(k_hexadecapole_gate<1>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, mu_hex, delta_mu, use_hex,
                               status, itmax);
            break;
        case 2:
            hipLaunchKernelGGL(k_hexadecapole_gate<2>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, mu_hex, delta_mu, use_hex,
                               status, itmax);
            break;
        default:
            hipLaunchKernelGGL(k_hexadecapole_gate<3>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, mu_hex, delta_mu, use_hex,
                               status, itmax);
            break;
    }
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}

extern "C" int caus_hexadecapole_vjp(void* stream, int32_t nlenses, const double* w, int64_t P,
                                     const double* params, int64_t params_stride,
                                     const double* rho, int64_t rho_stride, double u1,
                                     const double* g_mu, double* mu_hex, double* g_out,
                                     uint8_t* status, int32_t itmax) {
    if (P < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || params_stride < 0 || rho_stride < 0)
        return CAUS_ERR_ARG;
    if (P == 0) return CAUS_OK;
    if (!w || !rho || !g_mu || !mu_hex || !g_out || (nlenses > 1 && !params)) return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((P + kGateBlock - 1) / kGateBlock)), block(kGateBlock);
    const cd* wc = reinterpret_cast<const cd*>(w);
    switch (nlenses) {
        case 1:
            hipLaunchKernelGGL(k_hexadecapole_vjp<1>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, g_mu, mu_hex, g_out, status,
                               itmax);
            break;
        case 2:
            hipLaunchKernelGGL(k_hexadecapole_vjp<2>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, g_mu, mu_hex, g_out, status,
                               itmax);
            break;
        default:
            hipLaunchKernelGGL(k_hexadecapole_vjp<3>, grid, block, 0, s, wc, P, params,
                               params_stride, rho, rho_stride, u1, g_mu, mu_hex, g_out, status,
                               itmax);
            break;
    }
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}
