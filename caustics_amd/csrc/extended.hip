// C-ABI entry points of the extended-source contour kernels.
#include <atomic>

#include "extended_impl.hpp"

extern "C" int caus_mag_extended_source_vjp(void*, int32_t, const double*, int64_t, const double*,
                                            const double*, int64_t, const int32_t*, int32_t,
                                            int32_t, const double*, int64_t, int32_t,
                                            const double*, double*, double*, uint8_t*, void*,
                                            int64_t, int32_t, int32_t);
extern "C" int caus_mag_extended_source_work(void*, int32_t, const double*, int64_t, const double*,
                                             const double*, int64_t, const int32_t*, int32_t,
                                             int32_t, const double*, int64_t, int32_t,
                                             const double*, double*, double*, uint8_t*, void*,
                                             int64_t, int32_t, int32_t, uint64_t*);

using namespace caus;

extern "C" int64_t caus_extended_workspace_bytes(int32_t nlenses, int64_t S, int32_t npts_max) {
    if (nlenses < 1 || nlenses > 3 || S < 0 || npts_max < 4 || npts_max > 2 * kExtSortMax)
        return -1;
    const int deg = nlenses == 1 ? 2 : nlenses * nlenses + 1;
    return (int64_t)ext_total_bytes(deg, S, npts_max);
}

namespace caus {
// workgroups per CU of the flat (grid-stride) kernels; set by the launch-schedule tuner
static std::atomic<int> g_flat_per_cu{8};
}  // namespace caus

extern "C" int caus_ext_set_flat_blocks_per_cu(int32_t n) {
    if (n < 1 || n > 64) return CAUS_ERR_ARG;
    caus::g_flat_per_cu.store(n, std::memory_order_relaxed);
    return CAUS_OK;
}

namespace caus {
// grid of the flat extended-source kernels: enough lanes for the work list, at most
// g_flat_per_cu workgroups per CU (the kernels stride over their work list)
int ext_flat_grid(int64_t max_lanes) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            c > 0)
            cus = c;
    }
    const int64_t need = (max_lanes + kExtFlatBlock - 1) / kExtFlatBlock;
    const int64_t cap = (int64_t)cus * g_flat_per_cu.load(std::memory_order_relaxed);
    return (int)(need < 1 ? 1 : (need < cap ? need : cap));
}
}  // namespace caus

extern "C" int caus_mag_extended_source(void* stream, int32_t nlenses, const double* w0, int64_t S,
                                        const double* rho, const double* params,
                                        int64_t params_stride, const int32_t* npts_limb,
                                        int32_t npts_max, int32_t limb_darkening,
                                        const double* u1, int64_t u1_stride, int32_t npts_ld,
                                        double* mag, uint8_t* status, void* workspace,
                                        int64_t workspace_bytes, int32_t itmax, int32_t mode) {
    return caus_mag_extended_source_vjp(stream, nlenses, w0, S, rho, params, params_stride,
                                        npts_limb, npts_max, limb_darkening, u1, u1_stride,
                                        npts_ld, nullptr, mag, nullptr, status, workspace,
                                        workspace_bytes, itmax, mode);
}

extern "C" int caus_mag_extended_source_vjp(void* stream, int32_t nlenses, const double* w0,
                                            int64_t S, const double* rho, const double* params,
                                            int64_t params_stride, const int32_t* npts_limb,
                                            int32_t npts_max, int32_t limb_darkening,
                                            const double* u1, int64_t u1_stride, int32_t npts_ld,
                                            const double* g_mag, double* mag, double* g_params,
                                            uint8_t* status, void* workspace,
                                            int64_t workspace_bytes, int32_t itmax, int32_t mode) {
    return caus_mag_extended_source_work(stream, nlenses, w0, S, rho, params, params_stride,
                                         npts_limb, npts_max, limb_darkening, u1, u1_stride,
                                         npts_ld, g_mag, mag, g_params, status, workspace,
                                         workspace_bytes, itmax, mode, nullptr);
}

extern "C" int caus_mag_extended_source_work(void* stream, int32_t nlenses, const double* w0,
                                             int64_t S, const double* rho, const double* params,
                                             int64_t params_stride, const int32_t* npts_limb,
                                             int32_t npts_max, int32_t limb_darkening,
                                             const double* u1, int64_t u1_stride,
                                             int32_t npts_ld, const double* g_mag, double* mag,
                                             double* g_params, uint8_t* status, void* workspace,
                                             int64_t workspace_bytes, int32_t itmax, int32_t mode,
                                             uint64_t* work) {
    if (S < 0 || nlenses < 1 || nlenses > 3 || itmax < 0 || params_stride < 0 || u1_stride < 0)
        return CAUS_ERR_ARG;
    if (mode != CAUS_MODE_FP64 && mode != CAUS_MODE_FP64_COMP && mode != CAUS_MODE_CF32)
        return CAUS_ERR_ARG;
    if (npts_max < 4 || ext_mmax(npts_max) > kExtSortMax) return CAUS_ERR_ARG;
    if (limb_darkening && (npts_ld < 2 || npts_ld > 256 || !u1)) return CAUS_ERR_ARG;
    if (S == 0) return CAUS_OK;
    if (!w0 || !rho || !mag || !workspace || (nlenses > 1 && !params)) return CAUS_ERR_ARG;
    if (S > 0x7fffffff) return CAUS_ERR_ARG;
    const int deg = nlenses == 1 ? 2 : nlenses * nlenses + 1;
    const size_t per = ext_source_bytes(deg, ext_mmax(npts_max));
    if (workspace_bytes < (int64_t)ext_total_bytes(deg, S, npts_max)) return CAUS_ERR_ARG;
    ExtArgs a;
    a.w0 = reinterpret_cast<const cd*>(w0);
    a.S = S;
    a.rho = rho;
    a.params = params;
    a.pstride = params_stride;
    a.npts = npts_limb;
    a.npts_def = npts_max;
    a.u1 = u1;
    a.u1stride = u1_stride;
    a.npts_ld = npts_ld;
    a.mag = mag;
    a.status = status;
    a.g_mag = g_mag;
    a.g_out = g_params;
    const bool vjp = g_mag != nullptr;
    if (vjp && !g_params) return CAUS_ERR_ARG;
    a.ws = reinterpret_cast<char*>(workspace);
    a.work = reinterpret_cast<unsigned long long*>(work);
    a.ws_per_source = per;
    a.mmax = ext_mmax(npts_max);
    a.nadd_max = ext_n_add(npts_max);
    a.itmax = itmax;
    hipStream_t s = (hipStream_t)stream;
    switch (nlenses) {
        case 1:
            return mag_extended_nl<1>(s, a, mode, limb_darkening != 0, vjp);
        case 2:
            return mag_extended_nl<2>(s, a, mode, limb_darkening != 0, vjp);
        default:
            return mag_extended_nl<3>(s, a, mode, limb_darkening != 0, vjp);
    }
}

namespace {

// Greedy continuation of root sets along a sequence (match_points chained as in
// critical_and_caustic_curves, point_source.py:1628-1640): one lane per chain.
template <int DEG>
__global__ void k_match_chain(const cd* __restrict__ pts, cd* __restrict__ out, int64_t npts,
                              int64_t nchains) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchains) return;
    const cd* p = pts + c * npts * DEG;
    cd* o = out + c * npts * DEG;
    cd carry[DEG];
#pragma unroll
    for (int r = 0; r < DEG; ++r) carry[r] = p[r];
    for (int64_t k = 0; k < npts; ++k) {
        cd b[DEG];
#pragma unroll
        for (int r = 0; r < DEG; ++r) b[r] = p[k * DEG + r];
        const uint64_t perm = greedy_match<DEG>(carry, b);
        permute_pt<DEG>(b, perm);
#pragma unroll
        for (int r = 0; r < DEG; ++r) {
            carry[r] = b[r];
            o[k * DEG + r] = b[r];
        }
    }
}

}  // namespace

extern "C" int caus_match_chain(void* stream, const double* pts, double* out, int64_t npts,
                                int32_t n, int64_t nchains) {
    if (npts < 0 || nchains < 0 || n < 1) return CAUS_ERR_ARG;
    if (npts == 0 || nchains == 0) return CAUS_OK;
    if (!pts || !out || pts == out) return CAUS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const dim3 g((unsigned)((nchains + 63) / 64)), b(64);
    const cd* p = reinterpret_cast<const cd*>(pts);
    cd* o = reinterpret_cast<cd*>(out);
    switch (n) {
        case 2: hipLaunchKernelGGL(k_match_chain<2>, g, b, 0, s, p, o, npts, nchains); break;
        case 4: hipLaunchKernelGGL(k_match_chain<4>, g, b, 0, s, p, o, npts, nchains); break;
        case 5: hipLaunchKernelGGL(k_match_chain<5>, g, b, 0, s, p, o, npts, nchains); break;
        case 6: hipLaunchKernelGGL(k_match_chain<6>, g, b, 0, s, p, o, npts, nchains); break;
        case 10: hipLaunchKernelGGL(k_match_chain<10>, g, b, 0, s, p, o, npts, nchains); break;
        default: return CAUS_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}
