// C-ABI entry points of the fused point-source magnification and light-curve kernels.
#include <atomic>

#include "ps_impl.hpp"

using namespace caus;

// Scratch of caus_lightcurve_point_source_noalloc: the per-(light curve, epoch chunk)
// partial gradient sums that the caller's workspace holds in
// caus_lightcurve_point_source.  Every GPU the library runs on holds one copy until the
// process ends; calls on one GPU share it.
static constexpr int64_t kLcScratchBytes = CAUS_LC_NOALLOC_MAX_BYTES;
static __device__ __align__(256) double scratch[kLcScratchBytes / sizeof(double)];

namespace {

int np_for(int nl) { return nl == 1 ? 0 : (nl == 2 ? 2 : 5); }

int check_mode(int mode) {
    return (mode == CAUS_MODE_FP64 || mode == CAUS_MODE_FP64_COMP || mode == CAUS_MODE_CF32)
               ? CAUS_OK
               : CAUS_ERR_ARG;
}

int ps_dispatch(hipStream_t s, int nl, const cd* w, int64_t N, const double* params,
                int64_t pstride, const double* g_mag, double* mag, double* g_w, double* g_params,
                uint8_t* status, int itmax, int mode, bool vjp) {
    switch (nl) {
        case 1:
            return mag_point_source_nl<1>(s, w, N, params, pstride, g_mag, mag, g_w, g_params,
                                          status, itmax, mode, vjp);
        case 2:
            return mag_point_source_nl<2>(s, w, N, params, pstride, g_mag, mag, g_w, g_params,
                                          status, itmax, mode, vjp);
        case 3:
            return mag_point_source_nl<3>(s, w, N, params, pstride, g_mag, mag, g_w, g_params,
                                          status, itmax, mode, vjp);
        default:
            return CAUS_ERR_ARG;
    }
}

}  // namespace

extern "C" int caus_mag_point_source(void* stream, int32_t nlenses, const double* w, int64_t N,
                                     const double* params, int64_t params_stride, double* mag,
                                     uint8_t* status, int32_t itmax, int32_t mode) {
    if (N < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || check_mode(mode) != CAUS_OK)
        return CAUS_ERR_ARG;
    if (N == 0) return CAUS_OK;
    if (!w || !mag || (nlenses > 1 && !params) || params_stride < 0) return CAUS_ERR_ARG;
    return ps_dispatch((hipStream_t)stream, nlenses, reinterpret_cast<const cd*>(w), N, params,
                       params_stride, nullptr, mag, nullptr, nullptr, status, itmax, mode, false);
}

extern "C" int caus_mag_point_source_vjp(void* stream, int32_t nlenses, const double* w,
                                         int64_t N, const double* params, int64_t params_stride,
                                         const double* g_mag, double* mag, double* g_w,
                                         double* g_params, uint8_t* status, int32_t itmax,
                                         int32_t mode) {
    if (N < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || check_mode(mode) != CAUS_OK)
        return CAUS_ERR_ARG;
    if (N == 0) return CAUS_OK;
    if (!w || !mag || !g_mag || (nlenses > 1 && !params) || params_stride < 0)
        return CAUS_ERR_ARG;
    return ps_dispatch((hipStream_t)stream, nlenses, reinterpret_cast<const cd*>(w), N, params,
                       params_stride, g_mag, mag, g_w, g_params, status, itmax, mode, true);
}

extern "C" int64_t caus_lightcurve_workspace_bytes(int32_t nlenses, int64_t T, int64_t L) {
    if (nlenses < 1 || nlenses > 3 || T < 0 || L < 0) return -1;
    const int64_t per_block = (int64_t)kPsBlock * lc_epochs_per_lane(T, L);
    const int64_t nchunk = (T + per_block - 1) / per_block;
    return nchunk * L * (4 + np_for(nlenses)) * (int64_t)sizeof(double);
}

extern "C" int caus_lightcurve_point_source(void* stream, int32_t nlenses, const double* t,
                                            int64_t T, const double* lc_params, int64_t L,
                                            const double* g_mag, double* mag, double* g_params,
                                            double* workspace, int64_t workspace_bytes,
                                            uint8_t* status, int32_t itmax, int32_t mode) {
    if (T < 0 || L < 0 || itmax < 0 || nlenses < 1 || nlenses > 3 || check_mode(mode) != CAUS_OK)
        return CAUS_ERR_ARG;
    if (T == 0 || L == 0) return CAUS_OK;
    if (L > 65535 || !t || !lc_params || !mag) return CAUS_ERR_ARG;
    const bool vjp = g_mag != nullptr;
    if (vjp && (!g_params || !workspace ||
                workspace_bytes < caus_lightcurve_workspace_bytes(nlenses, T, L)))
        return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    switch (nlenses) {
        case 1:
            rc = lightcurve_point_source_nl<1>(s, t, T, lc_params, L, g_mag, mag, workspace,
                                               status, itmax, mode, vjp);
            break;
        case 2:
            rc = lightcurve_point_source_nl<2>(s, t, T, lc_params, L, g_mag, mag, workspace,
                                               status, itmax, mode, vjp);
            break;
        default:
            rc = lightcurve_point_source_nl<3>(s, t, T, lc_params, L, g_mag, mag, workspace,
                                               status, itmax, mode, vjp);
            break;
    }
    if (rc != CAUS_OK || !vjp) return rc;
    const int nq = 4 + np_for(nlenses);
    const int64_t per_block = (int64_t)kPsBlock * lc_epochs_per_lane(T, L);
    const int nchunk = (int)((T + per_block - 1) / per_block);
    const int64_t n = L * nq;
    hipLaunchKernelGGL(k_lightcurve_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       workspace, g_params, L, nchunk, nq);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}

namespace {
// device address of `scratch` on the current device, looked up once per device (the
// host-side name of a __device__ array is not a device address)
double* lc_scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    static std::atomic<double*> cached[64];
    double* p = cached[dev].load(std::memory_order_acquire);
    if (!p) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return nullptr;
        p = static_cast<double*>(q);
        cached[dev].store(p, std::memory_order_release);
    }
    return p;
}
}  // namespace

extern "C" int caus_lightcurve_point_source_noalloc(void* stream, int32_t nlenses,
                                                    const double* t, int64_t T,
                                                    const double* lc_params, int64_t L,
                                                    const double* g_mag, double* mag,
                                                    double* g_params, uint8_t* status,
                                                    int32_t itmax, int32_t mode) {
    if (nlenses < 1 || nlenses > 3 || T < 0 || L < 0) return CAUS_ERR_ARG;
    double* ws = nullptr;
    int64_t wsb = 0;
    if (g_mag && T > 0 && L > 0) {
        wsb = caus_lightcurve_workspace_bytes(nlenses, T, L);
        if (wsb < 0 || wsb > kLcScratchBytes) return CAUS_ERR_ARG;
        ws = lc_scratch_ptr();
        if (!ws) return CAUS_ERR_LAUNCH;
    }
    return caus_lightcurve_point_source(stream, nlenses, t, T, lc_params, L, g_mag, mag, g_params,
                                        ws, wsb, status, itmax, mode);
}
