// Source trajectory with annual parallax and its reverse mode (reference:
// AnnualParallaxTrajectory, src/caustics/trajectory.py:12-157).
//
// The Sun's position and velocity relative to the Earth, projected on the plane of the
// sky (east, north), come tabulated on a grid xp (the reference's daily t_jpl,
// trajectory.py:36, projected at :80-102).  Per curve (t0, u0, tE, pi_E, psi) and epoch t:
//   s(x)   = jax.numpy.interp(x, xp, s_tab)                  (:104-112)
//   ds     = s(t) - s(t0) - (t - t0) sdot(t0)                 (:114-116)
//   u_e    =  u0 cos psi + (t - t0)/tE sin psi + pi_E ds_e     (:149-152)
//   u_n    = -u0 sin psi + (t - t0)/tE cos psi + pi_E ds_n
//   w      = u_e + i u_n                                      (:155)
// with psi = atan2(pi_EE, pi_EN), pi_E = |(pi_EE, pi_EN)| in the cartesian
// parametrization (:140-142).  One workgroup per curve: its threads stride over the
// epochs, write w, and (reverse mode) sum the parameter cotangents in a fixed order
// (wave shuffles, then the four waves in order), so results are deterministic.  The
// derivative of interp in x is the slope of the segment that gives its value, zero
// outside the grid, as JAX differentiates it.
#include <hip/hip_runtime.h>

#include "caustics_hip.h"

namespace {

constexpr int kTrBlock = 256;

struct Seg {
    int64_t i;      // right end of the segment, 1..M-1
    int side;       // -1: x < xp[0], +1: x > xp[M-1], 0: inside (x may equal an end)
    double delta;   // x - xp[i-1]
    double dx;      // xp[i] - xp[i-1]
};

// jax.numpy.interp's segment: i = clip(#{k : xp[k] <= x}, 1, M-1).  The count is guessed
// from the mean grid step and corrected against xp itself, so any increasing grid
// works and a daily grid needs no correction step.
__device__ Seg find_segment(double x, const double* __restrict__ xp, int64_t M) {
    Seg s;
    const double x0 = xp[0], x1 = xp[M - 1];
    const double step = (x1 - x0) / (double)(M - 1);
    double g = (x - x0) / step + 1.0;
    g = fmin(fmax(g, 0.0), (double)M);  // also maps NaN to 0
    int64_t k = (int64_t)g;
    while (k < M && xp[k] <= x) ++k;
    while (k > 0 && xp[k - 1] > x) --k;
    s.i = k < 1 ? 1 : (k > M - 1 ? M - 1 : k);
    s.side = x < x0 ? -1 : (x > x1 ? 1 : 0);
    s.delta = x - xp[s.i - 1];
    s.dx = xp[s.i] - xp[s.i - 1];
    return s;
}

// value and slope of jax.numpy.interp(x, xp, fp) on segment s: fp[i-1] + (delta/dx) df,
// the end values outside the grid (slope 0 there); a segment of width at most 2^-104
// gives fp[i-1]
__device__ __forceinline__ void interp_at(const Seg& s, const double* __restrict__ fp,
                                          int64_t M, double& f, double& slope) {
    const double df = fp[s.i] - fp[s.i - 1];
    const bool dx0 = fabs(s.dx) <= 0x1p-104;  // numpy.spacing(float64 eps), as JAX
    f = dx0 ? fp[s.i - 1] : fp[s.i - 1] + (s.delta / s.dx) * df;
    slope = dx0 ? 0.0 : df / s.dx;
    if (s.side < 0) {
        f = fp[0];
        slope = 0.0;
    } else if (s.side > 0) {
        f = fp[M - 1];
        slope = 0.0;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// tab: (4, M) rows s_e, s_n, sdot_e, sdot_n; params (L, 5): t0, u0, tE and (pi_EE, pi_EN)
// or (pi_E, psi); w (L, T) complex; g_w (L, T) complex cotangents (torch convention:
// dL/dRe + i dL/dIm) or null; g_params (L, 5) out when g_w is given.
__global__ __launch_bounds__(kTrBlock) void k_parallax(
    const double* __restrict__ t, int64_t T, const double* __restrict__ xp,
    const double* __restrict__ tab, int64_t M, const double* __restrict__ params,
    int polar, double* __restrict__ w, const double* __restrict__ g_w,
    double* __restrict__ g_params) {
    __shared__ double red[kTrBlock / 64][5];
    const int64_t c = blockIdx.x;
    const double* prm = params + c * 5;
    const double t0 = prm[0], u0 = prm[1], tE = prm[2];
    double piE, psi;
    if (polar) {
        piE = prm[3];
        psi = prm[4];
    } else {
        psi = atan2(prm[3], prm[4]);
        piE = sqrt(prm[4] * prm[4] + prm[3] * prm[3]);
    }
    double sp, cp;
    sincos(psi, &sp, &cp);
    const double *se = tab, *sn = tab + M, *sde = tab + 2 * M, *sdn = tab + 3 * M;
    const Seg s0 = find_segment(t0, xp, M);
    double se0, sn0, sde0, sdn0, kse0, ksn0, ksde0, ksdn0;
    interp_at(s0, se, M, se0, kse0);
    interp_at(s0, sn, M, sn0, ksn0);
    interp_at(s0, sde, M, sde0, ksde0);
    interp_at(s0, sdn, M, sdn0, ksdn0);
    const double itE = 1.0 / tE;
    double a_t0 = 0.0, a_u0 = 0.0, a_tE = 0.0, a_piE = 0.0, a_psi = 0.0;
    for (int64_t e = threadIdx.x; e < T; e += kTrBlock) {
        const double te = t[e];
        const Seg st = find_segment(te, xp, M);
        double set, snt, k1, k2;
        interp_at(st, se, M, set, k1);
        interp_at(st, sn, M, snt, k2);
        const double dt = te - t0;
        const double dse = set - se0 - dt * sde0;
        const double dsn = snt - sn0 - dt * sdn0;
        const double tau = dt / tE;
        const double ue = u0 * cp + tau * sp + piE * dse;
        const double un = -u0 * sp + tau * cp + piE * dsn;
        if (w) {
            w[2 * (c * T + e)] = ue;
            w[2 * (c * T + e) + 1] = un;
        }
        if (g_w) {
            const double gr = g_w[2 * (c * T + e)], gi = g_w[2 * (c * T + e) + 1];
            // d ds / d t0 = -s'(t0) + sdot(t0) - (t - t0) sdot'(t0)
            const double ddse = -kse0 + sde0 - dt * ksde0;
            const double ddsn = -ksn0 + sdn0 - dt * ksdn0;
            a_t0 += gr * (piE * ddse - sp * itE) + gi * (piE * ddsn - cp * itE);
            a_u0 += gr * cp - gi * sp;
            a_tE -= (tau * itE) * (gr * sp + gi * cp);
            a_piE += gr * dse + gi * dsn;
            a_psi += gr * (tau * cp - u0 * sp) - gi * (u0 * cp + tau * sp);
        }
    }
    if (!g_w) return;
    double v[5] = {a_t0, a_u0, a_tE, a_piE, a_psi};
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        v[k] = wave_sum(v[k]);
        if (lane == 0) red[wv][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            r[k] = red[0][k];
#pragma unroll
            for (int q = 1; q < kTrBlock / 64; ++q) r[k] += red[q][k];
        }
        double* g = g_params + c * 5;
        g[0] = r[0];
        g[1] = r[1];
        g[2] = r[2];
        if (polar) {
            g[3] = r[3];
            g[4] = r[4];
        } else {
            // psi = atan2(pi_EE, pi_EN), pi_E = |(pi_EE, pi_EN)|; undefined at pi_E = 0
            const double pEE = prm[3], pEN = prm[4], ip2 = 1.0 / (piE * piE);
            g[3] = r[3] * (pEE / piE) + r[4] * (pEN * ip2);
            g[4] = r[3] * (pEN / piE) - r[4] * (pEE * ip2);
        }
    }
}

}  // namespace

extern "C" int caus_trajectory_annual_parallax(void* stream, const double* t, int64_t T,
                                               const double* xp, const double* sun_tab,
                                               int64_t M, const double* params, int64_t L,
                                               int32_t parametrization, double* w,
                                               const double* g_w, double* g_params) {
    if (T < 0 || L < 0 || M < 2 || (parametrization != 0 && parametrization != 1))
        return CAUS_ERR_ARG;
    if (L > 0x7fffffff || (g_w != nullptr) != (g_params != nullptr)) return CAUS_ERR_ARG;
    if (L == 0) return CAUS_OK;
    if (!xp || !sun_tab || !params || (T > 0 && !t) || (!w && !g_w)) return CAUS_ERR_ARG;
    hipLaunchKernelGGL(k_parallax, dim3((unsigned)L), dim3(kTrBlock), 0, (hipStream_t)stream, t,
                       T, xp, sun_tab, M, params, parametrization, w, g_w, g_params);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}
