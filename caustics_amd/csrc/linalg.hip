// Flux-marginalised log-likelihood of microlensing light curves with a diagonal data
// covariance (reference: marginalized_log_likelihood, src/caustics/linalg.py:10-70,
// the dense_covariance=False branch), for a batch of magnification curves that share
// one set of observations, and its gradient with respect to the magnifications.
//
// Per observatory segment the observed flux is modelled as f = F_s A + F_b.  With the
// design matrix M = [A, 1] and C^-1 = diag(c), the reference solves the normal
// equations N beta = M^T C^-1 f, N = M^T C^-1 M, Sigma = N^-1, and adds
//   -1/2 sum_i c_i (f_i - M_i beta)^2 + 1/2 log det(2 pi Sigma)
// over segments (linalg.py:59-69).  Here the same quantities come from weighted means
// and centred sums, which is the same algebra without forming the 2x2 system:
//   W = sum c, Abar = sum c A / W, fbar = sum c f / W, a = A - Abar, g = f - fbar,
//   Saa = sum c a^2, Sag = sum c a g,  F_s = Sag / Saa, F_b = fbar - F_s Abar,
//   det N = W Saa,  r = g - F_s a,  chi2 = sum c r^2,
//   ll = -chi2/2 + log(2 pi) - log(W Saa)/2.
// The gradient is the envelope derivative (beta is the least-squares optimum, so only
// the explicit dependence on A_i counts) plus the derivative of -log det N / 2:
//   d ll / d A_i = c_i (F_s r_i - a_i / Saa).
// One wavefront per curve walks its segments in order; every reduction is a fixed-order
// wave sum, so results are deterministic.  The magnifications of the first 1024 epochs of
// a segment are read from HBM once and held in registers for the three passes (later
// epochs are read again, mostly from cache); g_A is written once; f and c are shared by
// every curve and stay in cache.
#include <hip/hip_runtime.h>

#include "caustics_hip.h"

namespace {

constexpr int kLLBlock = 256;               // 4 waves, one curve per wave
constexpr int kLLCurves = kLLBlock / 64;
constexpr int kLLCache = 16;                // epochs of A per lane held in registers
constexpr double kTwoPi = 6.28318530717958647692;

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kLLBlock) void k_marginalized_ll(
    const double* __restrict__ A, int64_t B, int64_t n, int64_t lda,
    const double* __restrict__ f, const double* __restrict__ c, const int64_t* __restrict__ off,
    int nseg, const double* __restrict__ g_ll, double* __restrict__ ll,
    double* __restrict__ beta, double* __restrict__ gA, int64_t ldg) {
    const int lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * kLLCurves + (threadIdx.x >> 6);
    if (b >= B) return;  // whole waves leave; nothing below synchronises the block
    const double* a = A + b * lda;
    double* ga = gA ? gA + b * ldg : nullptr;
    const double gb = gA ? (g_ll ? g_ll[b] : 1.0) : 0.0;
    double total = 0.0;
    for (int k = 0; k < nseg; ++k) {
        // segment bounds, clamped so that malformed offsets cannot leave [0, n)
        const int64_t i0 = min(max(off[k], (int64_t)0), n);
        const int64_t i1 = min(max(off[k + 1], i0), n);
        // the first 64 x kLLCache magnifications of the segment are read from HBM once,
        // every load of a lane issued together, and kept in registers for the three
        // passes; f and c are shared by all curves and come from cache; an epoch
        // outside the segment has weight 0 and adds nothing
        double ra[kLLCache];
#pragma unroll
        for (int j = 0; j < kLLCache; ++j) {
            const int64_t i = i0 + lane + 64 * j;
            ra[j] = i < i1 ? a[i] : 0.0;
        }
        const int64_t i_tail = i0 + lane + 64 * kLLCache;
        double W = 0.0, SA = 0.0, SF = 0.0;
#pragma unroll
        for (int j = 0; j < kLLCache; ++j) {
            const int64_t i = i0 + lane + 64 * j;
            const double ci = i < i1 ? c[i] : 0.0, fi = i < i1 ? f[i] : 0.0;
            W += ci;
            SA = fma(ci, ra[j], SA);
            SF = fma(ci, fi, SF);
        }
        for (int64_t i = i_tail; i < i1; i += 64) {
            const double ci = c[i];
            W += ci;
            SA = fma(ci, a[i], SA);
            SF = fma(ci, f[i], SF);
        }
        W = wave_sum64(W);
        SA = wave_sum64(SA);
        SF = wave_sum64(SF);
        const double Abar = SA / W, fbar = SF / W;
        double Saa = 0.0, Sag = 0.0;
#pragma unroll
        for (int j = 0; j < kLLCache; ++j) {
            const int64_t i = i0 + lane + 64 * j;
            const double ci = i < i1 ? c[i] : 0.0, fi = i < i1 ? f[i] : 0.0;
            const double da = ra[j] - Abar, dg = fi - fbar;
            const double cda = ci * da;
            Saa = fma(cda, da, Saa);
            Sag = fma(cda, dg, Sag);
        }
        for (int64_t i = i_tail; i < i1; i += 64) {
            const double da = a[i] - Abar, dg = f[i] - fbar;
            const double cda = c[i] * da;
            Saa = fma(cda, da, Saa);
            Sag = fma(cda, dg, Sag);
        }
        Saa = wave_sum64(Saa);
        Sag = wave_sum64(Sag);
        const double Fs = Sag / Saa;
        const double Fb = fbar - Fs * Abar;
        const double iSaa = 1.0 / Saa;
        double chi2 = 0.0;
#pragma unroll
        for (int j = 0; j < kLLCache; ++j) {
            const int64_t i = i0 + lane + 64 * j;
            if (i < i1) {
                const double ci = c[i];
                const double da = ra[j] - Abar;
                const double r = (f[i] - fbar) - Fs * da;
                chi2 = fma(ci * r, r, chi2);
                if (ga) ga[i] = gb * ci * (Fs * r - da * iSaa);
            }
        }
        for (int64_t i = i_tail; i < i1; i += 64) {
            const double ci = c[i];
            const double da = a[i] - Abar;
            const double r = (f[i] - fbar) - Fs * da;
            chi2 = fma(ci * r, r, chi2);
            if (ga) ga[i] = gb * ci * (Fs * r - da * iSaa);
        }
        chi2 = wave_sum64(chi2);
        total += -0.5 * chi2 + (log(kTwoPi) - 0.5 * log(W * Saa));
        if (beta && lane == 0) {
            beta[(b * nseg + k) * 2] = Fs;
            beta[(b * nseg + k) * 2 + 1] = Fb;
        }
    }
    if (lane == 0) ll[b] = total;
}

}  // namespace

extern "C" int caus_marginalized_log_likelihood(void* stream, const double* A, int64_t B,
                                                int64_t n, int64_t lda, const double* fobs,
                                                const double* c_inv, const int64_t* seg_off,
                                                int32_t nseg, double* ll, double* beta,
                                                const double* g_ll, double* g_A, int64_t ldg) {
    if (B < 0 || n < 0 || nseg < 1 || lda < n || (g_A && ldg < n)) return CAUS_ERR_ARG;
    if (B == 0) return CAUS_OK;
    if (B > (int64_t(1) << 26) || !A || !fobs || !c_inv || !seg_off || !ll) return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = (unsigned)((B + kLLCurves - 1) / kLLCurves);
    hipLaunchKernelGGL(k_marginalized_ll, dim3(nb), dim3(kLLBlock), 0, s, A, B, n, lda, fobs,
                       c_inv, seg_off, (int)nseg, g_ll, ll, beta, g_A, ldg);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}
