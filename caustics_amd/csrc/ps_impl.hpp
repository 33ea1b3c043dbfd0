// Fused point-source magnification kernels (forward and forward+VJP) and the fused
// rectilinear light-curve kernel, instantiated per lens count in ps_nlN.hip.
//
// One evaluation per lane: lens frame and coefficients in FP64 registers, all-roots
// solve, lens-equation filter (|lens_eq(z) - w| < 1e-6, point_source.py:1703-1707),
// sum of 1/|det J| over real images (point_source.py:1828-1830).  The VJP re-solves
// nothing: one solve feeds the value and the implicit-derivative tangents of every
// parameter (forward-mode tangents inside the lane, reverse over the batch).
#pragma once
#include "caustics_hip.h"
#include "cf32.hpp"
#include "lens.hpp"
#include "solver.hpp"

namespace caus {

constexpr int kPsBlock = 256;

template <int NL>
constexpr int ps_degree() {
    return NL * NL + 1;
}

// warm = true: z holds initial estimates (the roots of a nearby source position)
// POLISH: one compensated Newton step per root after convergence (polish_roots).
template <int NL, int MODE, bool POLISH = false>
__device__ __forceinline__ bool solve_lens(cd wl, const LensFrame<NL>& lf, cd (&z)[ps_degree<NL>()],
                                           int itmax, bool warm = false,
                                           SolveStats* stats = nullptr) {
    constexpr int DEG = ps_degree<NL>();
    cd p[DEG + 1];
    lens_poly<NL>(wl, lf, p);
    SolveStats st;
    bool ok;
    if constexpr (MODE == MODE_CF32)
        ok = solve_cf32<DEG>(p, z, itmax, warm, st);
    else
        ok = solve_fp64<DEG, MODE, false>(p, z, itmax, warm, st);
    if constexpr (POLISH) polish_roots<DEG>(p, z);
    if (stats) {
        stats->sweeps += st.sweeps;
        stats->it64 += st.it64;
        stats->itcomp += st.itcomp;
        stats->it32 += st.it32;
    }
    return ok;
}

// The reference-compatible FP64 solve (reference initial estimates and branch structure:
// the reference solver's root order) from scratch, then the compensated polish.
template <int NL>
__device__ __forceinline__ bool solve_lens_refcompat(cd wl, const LensFrame<NL>& lf,
                                                     cd (&z)[ps_degree<NL>()], int itmax,
                                                     SolveStats* stats = nullptr) {
    constexpr int DEG = ps_degree<NL>();
    cd p[DEG + 1];
    lens_poly<NL>(wl, lf, p);
    SolveStats st;
    const bool ok = solve_fp64<DEG, MODE_FP64, true>(p, z, itmax, false, st);
    polish_roots<DEG>(p, z);
    if (stats) {
        stats->sweeps += st.sweeps;
        stats->it64 += st.it64;
    }
    return ok;
}

// Value and tangents of the magnification at lens-frame source position wl.
// Tangent direction d has source-plane tangent dw[d] and lens-frame tangent lt[d].
template <int NL, int MODE, int ND>
__device__ __forceinline__ double mag_with_tangents(cd wl, const LensFrame<NL>& lf,
                                                    const cd (&dw)[ND > 0 ? ND : 1],
                                                    const LensTangent<NL> (&lt)[ND > 0 ? ND : 1],
                                                    double (&dm)[ND > 0 ? ND : 1], int itmax,
                                                    bool& ok, cd* zio = nullptr,
                                                    bool warm = false) {
#pragma unroll
    for (int d = 0; d < ND; ++d) dm[d] = 0.0;
    double m = 0.0;
    if constexpr (NL == 1) {
        // closed form, point_source.py:1664-1671; det = 1 - 1/|z|^4 (:1562-1563)
        const double wa = norm2(wl);
        const double sq = sqrt(1.0 + 4.0 * rcp(wa));
        cd z[2] = {(0.5 * (1.0 + sq)) * wl, (0.5 * (1.0 - sq)) * wl};
        ok = true;
        LensFrame<1> f1 = lf;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            ImageTerms<1> it = image_terms<1>(z[j], f1);
            const double det = 1.0 - norm2(it.K);
            m += fabs(rcp(det));
#pragma unroll
            for (int d = 0; d < ND; ++d) dm[d] += dmu_image<1>(it, f1, det, dw[d], lt[d]);
        }
        return m;
    } else {
        constexpr int DEG = ps_degree<NL>();
        cd z[DEG];
        if (warm) {
#pragma unroll
            for (int j = 0; j < DEG; ++j) z[j] = zio[j];
        }
        ok = solve_lens<NL, MODE>(wl, lf, z, itmax, warm);
        if (zio) {
#pragma unroll
            for (int j = 0; j < DEG; ++j) zio[j] = z[j];
        }
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            ImageTerms<NL> it = image_terms<NL>(z[j], lf);
            const cd res = z[j] - it.G - wl;
            if (norm2(res) < 1e-12) {
                const double det = 1.0 - norm2(it.K);
                m += fabs(rcp(det));
#pragma unroll
                for (int d = 0; d < ND; ++d) dm[d] += dmu_image<NL>(it, lf, det, dw[d], lt[d]);
            }
        }
        return m;
    }
}

// ---------------------------------------------------------------------------
// point-source magnification at arbitrary source positions
//   VJP=false: mag only.  VJP=true: also g_w = g*(dm/dRe w + i dm/dIm w) and
//   g_params[i][k] = g * dm/dparam_k (per evaluation).
template <int NL, int MODE, bool VJP>
__global__ __launch_bounds__(kPsBlock) void k_mag_point_source(
    const cd* __restrict__ w, int64_t N, const double* __restrict__ params, int64_t pstride,
    const double* __restrict__ g_mag, double* __restrict__ mag, double* __restrict__ g_w,
    double* __restrict__ g_params, uint8_t* __restrict__ status, int itmax) {
    constexpr int NP = NL == 1 ? 0 : n_user_params<NL>();
    constexpr int ND = VJP ? 2 + NP : 0;
    const int64_t i = (int64_t)blockIdx.x * kPsBlock + threadIdx.x;
    if (i >= N) return;
    const double* prm = params + i * pstride;
    LensFrame<NL> lf;
    double x_cm = 0.0;
    if constexpr (NL == 1) {
        lf.e[0] = 1.0;
        lf.r[0] = cd{0.0, 0.0};
    } else {
        lens_frame<NL>(prm, lf, x_cm);
    }
    const cd wl = w[i] + cd{x_cm, 0.0};
    cd dw[ND > 0 ? ND : 1];
    LensTangent<NL> lt[ND > 0 ? ND : 1];
    double dm[ND > 0 ? ND : 1];
    if constexpr (VJP) {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                lt[d].de[j] = 0.0;
                lt[d].dr[j] = cd{0.0, 0.0};
            }
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
    }
    bool ok;
    const double m = mag_with_tangents<NL, MODE, ND>(wl, lf, dw, lt, dm, itmax, ok);
    mag[i] = m;
    if (status) status[i] = ok ? 1 : 0;
    if constexpr (VJP) {
        const double g = g_mag[i];
        if (g_w) {
            g_w[2 * i] = g * dm[0];
            g_w[2 * i + 1] = g * dm[1];
        }
        if (g_params) {
#pragma unroll
            for (int k = 0; k < NP; ++k) g_params[i * NP + k] = g * dm[2 + k];
        }
    }
}

// ---------------------------------------------------------------------------
// Light curves: L parameter sets x T epochs.  Parameter row l: t0, u0, tE, alpha,
// then the lens parameters.  w(t) = e^{-i alpha}(u0 + i tau), tau = (t - t0)/tE
// (trajectory.py:152-155 without parallax, psi -> alpha).  Grid (ceil(T/256), L):
// each block is one light curve's chunk of epochs; with VJP the block reduces
// g_mag * dm/dparam over its epochs in a fixed order into ws[l][chunk][param].
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// epochs handled by one lane (consecutive, each warm-started from the previous one's
// roots): enough lanes to fill the device at two waves per SIMD, at most 8 per lane
__host__ __device__ inline int lc_epochs_per_lane(int64_t T, int64_t L) {
    const int64_t e = (T * L + 131071) / 131072;
    return (int)(e < 1 ? 1 : (e > 8 ? 8 : e));
}

template <int NL, int MODE, bool VJP>
__global__ __launch_bounds__(kPsBlock) void k_lightcurve_point_source(
    const double* __restrict__ t, int64_t T, const double* __restrict__ lcp,
    const double* __restrict__ g_mag, double* __restrict__ mag, double* __restrict__ ws,
    uint8_t* __restrict__ status, int itmax, int epl) {
    constexpr int NP = NL == 1 ? 0 : n_user_params<NL>();
    constexpr int NQ = 4 + NP;
    constexpr int ND = VJP ? NQ : 0;
    constexpr int DEG = NL == 1 ? 2 : ps_degree<NL>();
    __shared__ double red[kPsBlock / 64][NQ];
    const int64_t l = blockIdx.y;
    const double* prm = lcp + l * NQ;
    double gsum[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) gsum[k] = 0.0;
    cd zw[DEG];
    bool warm = false;
    const int64_t ti0 = ((int64_t)blockIdx.x * kPsBlock + threadIdx.x) * epl;
    for (int ep = 0; ep < epl; ++ep) {
    const int64_t ti = ti0 + ep;
    const bool active = ti < T;
    const double t0 = prm[0], u0 = prm[1], tE = prm[2], alpha = prm[3];
    LensFrame<NL> lf;
    double x_cm = 0.0;
    if constexpr (NL == 1) {
        lf.e[0] = 1.0;
        lf.r[0] = cd{0.0, 0.0};
    } else {
        lens_frame<NL>(prm + 4, lf, x_cm);
    }
    const double tt = active ? t[ti] : t0;
    const double itE = rcp(tE);
    const double tau = (tt - t0) * itE;
    double sa, ca;
    sincos(alpha, &sa, &ca);
    const cd ea = cd{ca, -sa};                 // e^{-i alpha}
    const cd w = ea * cd{u0, tau};
    cd dw[ND > 0 ? ND : 1];
    LensTangent<NL> lt[ND > 0 ? ND : 1];
    double dm[ND > 0 ? ND : 1];
    if constexpr (VJP) {
#pragma unroll
        for (int d = 0; d < 4; ++d) {
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                lt[d].de[j] = 0.0;
                lt[d].dr[j] = cd{0.0, 0.0};
            }
        }
        const cd iea = cd{-ea.im, ea.re};      // i e^{-i alpha} = dw/dtau
        dw[0] = (-itE) * iea;                  // t0
        dw[1] = ea;                            // u0
        dw[2] = (-tau * itE) * iea;            // tE
        dw[3] = cd{w.im, -w.re};               // alpha: -i w
        if constexpr (NP > 0) {
            LensTangent<NL> up[NP];
            lens_frame_tangents<NL>(prm + 4, up);
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                lt[4 + k] = up[k];
                dw[4 + k] = cd{up[k].dxcm, 0.0};
            }
        }
    }
    bool ok = true;
    double m = 0.0;
    if (active) {
        m = mag_with_tangents<NL, MODE, ND>(w + cd{x_cm, 0.0}, lf, dw, lt, dm, itmax, ok, zw, warm);
        warm = ok;
        mag[l * T + ti] = m;
        if (status) status[l * T + ti] = ok ? 1 : 0;
        if constexpr (VJP) {
            const double g = g_mag[l * T + ti];
#pragma unroll
            for (int k = 0; k < NQ; ++k) gsum[k] += g * dm[k];
        }
    }
    }
    if constexpr (VJP) {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const double v = wave_sum(gsum[k]);
            if (lane == 0) red[wv][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < NQ) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < kPsBlock / 64; ++q) s += red[q][threadIdx.x];
            ws[(l * gridDim.x + blockIdx.x) * NQ + threadIdx.x] = s;
        }
    }
}

// g_params[l][k] = sum over chunks c (in order) of ws[l][c][k]
static __global__ void k_lightcurve_finalize(const double* __restrict__ ws, double* __restrict__ g,
                                      int64_t L, int nchunk, int nq) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L * nq) return;
    const int64_t l = i / nq, k = i % nq;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += ws[(l * nchunk + c) * nq + k];
    g[i] = s;
}

template <int NL>
int mag_point_source_nl(hipStream_t s, const cd* w, int64_t N, const double* params,
                        int64_t pstride, const double* g_mag, double* mag, double* g_w,
                        double* g_params, uint8_t* status, int itmax, int mode, bool vjp);
template <int NL>
int lightcurve_point_source_nl(hipStream_t s, const double* t, int64_t T, const double* lcp,
                               int64_t L, const double* g_mag, double* mag, double* ws,
                               uint8_t* status, int itmax, int mode, bool vjp);

#define CAUS_DEFINE_PS(NL)                                                                       \
    template <>                                                                                  \
    int mag_point_source_nl<NL>(hipStream_t s, const cd* w, int64_t N, const double* params,    \
                                int64_t pstride, const double* g_mag, double* mag, double* g_w, \
                                double* g_params, uint8_t* status, int itmax, int mode,          \
                                bool vjp) {                                                      \
        const dim3 grid((unsigned)((N + kPsBlock - 1) / kPsBlock)), block(kPsBlock);             \
        switch (mode * 2 + (vjp ? 1 : 0)) {                                                      \
            case 0:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_FP64, false>), grid, block, 0, s, \
                                   w, N, params, pstride, g_mag, mag, g_w, g_params, status,     \
                                   itmax);                                                       \
                break;                                                                           \
            case 1:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_FP64, true>), grid, block, 0, s,  \
                                   w, N, params, pstride, g_mag, mag, g_w, g_params, status,     \
                                   itmax);                                                       \
                break;                                                                           \
            case 2:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_FP64_COMP, false>), grid, block, \
                                   0, s, w, N, params, pstride, g_mag, mag, g_w, g_params,       \
                                   status, itmax);                                               \
                break;                                                                           \
            case 3:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_FP64_COMP, true>), grid, block,  \
                                   0, s, w, N, params, pstride, g_mag, mag, g_w, g_params,       \
                                   status, itmax);                                               \
                break;                                                                           \
            case 4:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_CF32, false>), grid, block, 0, s, \
                                   w, N, params, pstride, g_mag, mag, g_w, g_params, status,     \
                                   itmax);                                                       \
                break;                                                                           \
            case 5:                                                                              \
                hipLaunchKernelGGL((k_mag_point_source<NL, MODE_CF32, true>), grid, block, 0, s,  \
                                   w, N, params, pstride, g_mag, mag, g_w, g_params, status,     \
                                   itmax);                                                       \
                break;                                                                           \
            default:                                                                             \
                return CAUS_ERR_ARG;                                                             \
        }                                                                                        \
        return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;                      \
    }                                                                                            \
    template <>                                                                                  \
    int lightcurve_point_source_nl<NL>(hipStream_t s, const double* t, int64_t T,               \
                                       const double* lcp, int64_t L, const double* g_mag,        \
                                       double* mag, double* ws, uint8_t* status, int itmax,      \
                                       int mode, bool vjp) {                                     \
        const int epl = lc_epochs_per_lane(T, L);                                                \
        const int64_t per_block = (int64_t)kPsBlock * epl;                                       \
        const dim3 grid((unsigned)((T + per_block - 1) / per_block), (unsigned)L), block(kPsBlock); \
        switch (mode * 2 + (vjp ? 1 : 0)) {                                                      \
            case 0:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_FP64, false>), grid,      \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            case 1:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_FP64, true>), grid,       \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            case 2:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_FP64_COMP, false>), grid, \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            case 3:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_FP64_COMP, true>), grid,  \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            case 4:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_CF32, false>), grid,      \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            case 5:                                                                              \
                hipLaunchKernelGGL((k_lightcurve_point_source<NL, MODE_CF32, true>), grid,       \
                                   block, 0, s, t, T, lcp, g_mag, mag, ws, status, itmax, epl);       \
                break;                                                                           \
            default:                                                                             \
                return CAUS_ERR_ARG;                                                             \
        }                                                                                        \
        return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;                      \
    }

}  // namespace caus
