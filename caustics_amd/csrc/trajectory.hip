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
// parametrization (:140-142).  One workgroup per kTrSets parameter sets: its threads
// stride over the epochs, write w, and (reverse mode) sum the parameter cotangents in a
// fixed order (wave shuffles, then the four waves in order), so results are
// deterministic.  The
// derivative of interp in x is the slope of the segment that gives its value, zero
// outside the grid, as JAX differentiates it.
#include <hip/hip_runtime.h>

#include "caustics_hip.h"

namespace {

constexpr int kTrBlock = 256;
constexpr int kTrSets = 8;    // most parameter sets per workgroup
constexpr int kTrEpochs = 8;  // epochs per thread and chunk

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

// Per parameter set: the parameters, the trigonometry of psi and the interpolants at t0
// with their slopes (the t0 dependence of s(t0) and sdot(t0)).
struct CurveK {
    double t0, u0, tE, itE, piE, sp, cp, pEE, pEN;
    double se0, sn0, sde0, sdn0, kse0, ksn0, ksde0, ksdn0;
};

// tab: (4, M) rows s_e, s_n, sdot_e, sdot_n; params (L, 5): t0, u0, tE and (pi_EE, pi_EN)
// or (pi_E, psi); w (L, T) complex; g_w (L, T) complex cotangents (torch convention:
// dL/dRe + i dL/dIm); g_params (L, 5) out (reverse mode).  A workgroup holds `sets`
// (<= kTrSets) parameter sets and takes the epochs in chunks of kTrBlock * kTrEpochs: each thread
// locates the grid segments of its kTrEpochs epochs once and evaluates every set of the
// workgroup there.
template <bool VJP>
__global__ __launch_bounds__(kTrBlock) void k_parallax(
    const double* __restrict__ t, int64_t T, const double* __restrict__ xp,
    const double* __restrict__ tab, int64_t M, const double* __restrict__ params, int64_t L,
    int sets, int polar, double* __restrict__ w, const double* __restrict__ g_w,
    double* __restrict__ g_params) {
    __shared__ CurveK ck[kTrSets];
    __shared__ double red[kTrBlock / 64][kTrSets][5];
    const int64_t c0 = (int64_t)blockIdx.x * sets;
    const int nset = (int)(L - c0 < sets ? L - c0 : sets);
    const double *se = tab, *sn = tab + M, *sde = tab + 2 * M, *sdn = tab + 3 * M;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < nset) {
        const double* prm = params + (c0 + threadIdx.x) * 5;
        CurveK k;
        k.t0 = prm[0];
        k.u0 = prm[1];
        k.tE = prm[2];
        k.itE = 1.0 / k.tE;
        double psi;
        k.pEE = prm[3];
        k.pEN = prm[4];
        if (polar) {
            k.piE = prm[3];
            psi = prm[4];
        } else {
            psi = atan2(prm[3], prm[4]);
            k.piE = sqrt(prm[4] * prm[4] + prm[3] * prm[3]);
        }
        sincos(psi, &k.sp, &k.cp);
        const Seg s0 = find_segment(k.t0, xp, M);
        interp_at(s0, se, M, k.se0, k.kse0);
        interp_at(s0, sn, M, k.sn0, k.ksn0);
        interp_at(s0, sde, M, k.sde0, k.ksde0);
        interp_at(s0, sdn, M, k.sdn0, k.ksdn0);
        ck[threadIdx.x] = k;
    }
    if (VJP && threadIdx.x < (kTrBlock / 64) * kTrSets * 5) (&red[0][0][0])[threadIdx.x] = 0.0;
    __syncthreads();
    for (int64_t e0 = 0; e0 < T; e0 += (int64_t)kTrBlock * kTrEpochs) {
        double te[kTrEpochs], set[kTrEpochs], snt[kTrEpochs];
#pragma unroll
        for (int j = 0; j < kTrEpochs; ++j) {
            const int64_t e = e0 + threadIdx.x + (int64_t)j * kTrBlock;
            te[j] = e < T ? t[e] : 0.0;
            double k1, k2;
            const Seg st = find_segment(te[j], xp, M);
            interp_at(st, se, M, set[j], k1);
            interp_at(st, sn, M, snt[j], k2);
        }
        for (int c = 0; c < nset; ++c) {
            const CurveK k = ck[c];
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
#pragma unroll
            for (int j = 0; j < kTrEpochs; ++j) {
                const int64_t e = e0 + threadIdx.x + (int64_t)j * kTrBlock;
                if (e >= T) break;
                const double dt = te[j] - k.t0;
                const double dse = set[j] - k.se0 - dt * k.sde0;
                const double dsn = snt[j] - k.sn0 - dt * k.sdn0;
                const double tau = dt / k.tE;
                const int64_t o = 2 * ((c0 + c) * T + e);
                if (!VJP) {
                    w[o] = k.u0 * k.cp + tau * k.sp + k.piE * dse;
                    w[o + 1] = -k.u0 * k.sp + tau * k.cp + k.piE * dsn;
                } else {
                    const double gr = g_w[o], gi = g_w[o + 1];
                    // d ds / d t0 = -s'(t0) + sdot(t0) - (t - t0) sdot'(t0)
                    const double ddse = -k.kse0 + k.sde0 - dt * k.ksde0;
                    const double ddsn = -k.ksn0 + k.sdn0 - dt * k.ksdn0;
                    a0 += gr * (k.piE * ddse - k.sp * k.itE) + gi * (k.piE * ddsn - k.cp * k.itE);
                    a1 += gr * k.cp - gi * k.sp;
                    a2 -= (tau * k.itE) * (gr * k.sp + gi * k.cp);
                    a3 += gr * dse + gi * dsn;
                    a4 += gr * (tau * k.cp - k.u0 * k.sp) - gi * (k.u0 * k.cp + tau * k.sp);
                }
            }
            if (VJP) {
                a0 = wave_sum(a0);
                a1 = wave_sum(a1);
                a2 = wave_sum(a2);
                a3 = wave_sum(a3);
                a4 = wave_sum(a4);
                if (lane == 0) {
                    red[wv][c][0] += a0;
                    red[wv][c][1] += a1;
                    red[wv][c][2] += a2;
                    red[wv][c][3] += a3;
                    red[wv][c][4] += a4;
                }
            }
        }
    }
    if (!VJP) return;
    __syncthreads();
    if (threadIdx.x < nset) {
        const int c = threadIdx.x;
        double r[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            r[q] = red[0][c][q];
#pragma unroll
            for (int u = 1; u < kTrBlock / 64; ++u) r[q] += red[u][c][q];
        }
        const CurveK& k = ck[c];
        double* g = g_params + (c0 + c) * 5;
        g[0] = r[0];
        g[1] = r[1];
        g[2] = r[2];
        if (polar) {
            g[3] = r[3];
            g[4] = r[4];
        } else {
            // psi = atan2(pi_EE, pi_EN), pi_E = |(pi_EE, pi_EN)|; undefined at pi_E = 0
            const double ip2 = 1.0 / (k.piE * k.piE);
            g[3] = r[3] * (k.pEE / k.piE) + r[4] * (k.pEN * ip2);
            g[4] = r[3] * (k.pEN / k.piE) - r[4] * (k.pEE * ip2);
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
    // sets per workgroup: as many as keep at least 2048 workgroups (8 per compute unit)
    const int sets = (int)(L / 2048 < 1 ? 1 : (L / 2048 > kTrSets ? kTrSets : L / 2048));
    const unsigned blocks = (unsigned)((L + sets - 1) / sets);
    hipStream_t s = (hipStream_t)stream;
    if (w)
        hipLaunchKernelGGL(k_parallax<false>, dim3(blocks), dim3(kTrBlock), 0, s, t, T, xp,
                           sun_tab, M, params, L, sets, parametrization, w, g_w, g_params);
    if (g_w)
        hipLaunchKernelGGL(k_parallax<true>, dim3(blocks), dim3(kTrBlock), 0, s, t, T, xp,
                           sun_tab, M, params, L, sets, parametrization, w, g_w, g_params);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}
