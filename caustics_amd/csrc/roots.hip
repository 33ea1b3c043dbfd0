// Batched all-roots solve: kernels and C-ABI entry points.
//
// Replaces the reference's XLA custom calls cpu_ehrlich_aberth (lib/ehrlich_aberth/
// cpu_ops.cc:15-81) and gpu_ehrlich_aberth (lib/ehrlich_aberth/kernels.cc.cu:52-104):
// coefficients (N, deg+1) complex128 ascending, optional initial roots (N, deg),
// output roots (N, deg).  One polynomial per lane, 256 lanes per workgroup, no
// allocation, no device synchronisation.  Scratch memory: the reference-order (RefCompat)
// initial estimates read their log-magnitude table at runtime hull vertices, (deg+1)
// doubles per lane; the Fast form's estimates and sweeps stay in registers.
#include <atomic>

#include "caustics_hip.h"
#include "roots_impl.hpp"

using namespace caus;

namespace caus {
// Workgroups of the lane-refill root kernels (k_roots_persistent: reference-compatible
// and compensated forms, and the Fast form below degree 5): 8 per compute unit of the
// current device.  That is more than fit at once (3-5 per compute unit, by degree and
// form); the surplus starts as resident workgroups finish and takes its static share of
// the batch, which measured a few percent faster than a grid of exactly the resident
// count.  The compacted Fast-form kernel runs one polynomial per lane instead.
int roots_grid_limit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    // one value per device; concurrent first calls compute the same value
    static std::atomic<int> cached[64];
    if (dev < 0 || dev >= 64) return 1024;
    int v = cached[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        v = 8 * cus;
        cached[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}
}  // namespace caus


extern "C" int caus_roots(void* stream, const double* coeffs, const double* roots_init,
                          double* roots, int32_t* stats, uint8_t* status, int64_t N, int32_t deg,
                          int32_t itmax, int32_t mode, int32_t init_mode) {
    if (N < 0 || itmax < 0 || (N > 0 && (!coeffs || !roots))) return CAUS_ERR_ARG;
    if (init_mode != CAUS_INIT_REFERENCE && init_mode != CAUS_INIT_COMPLEX) return CAUS_ERR_ARG;
    if (N == 0) return CAUS_OK;
    hipStream_t s = (hipStream_t)stream;
    const cd* c = reinterpret_cast<const cd*>(coeffs);
    const cd* ri = reinterpret_cast<const cd*>(roots_init);
    cd* r = reinterpret_cast<cd*>(roots);
    switch (deg) {
#define CAUS_DEG(D) \
    case D:         \
        return roots_for_degree<D>(s, c, ri, r, stats, status, N, itmax, mode, init_mode);
        CAUS_DEG(1)
        CAUS_DEG(2)
        CAUS_DEG(3)
        CAUS_DEG(4)
        CAUS_DEG(5)
        CAUS_DEG(6)
        CAUS_DEG(7)
        CAUS_DEG(8)
        CAUS_DEG(9)
        CAUS_DEG(10)
        CAUS_DEG(11)
        CAUS_DEG(12)
#undef CAUS_DEG
        default:
            return CAUS_ERR_ARG;
    }
}

// Reference-layout GPU target (kernels.h:11-20, gpu_ops.cc:14-26): buffers = {coeffs,
// roots}, opaque = EhrlichAberthDescriptor bytes.  Unlike the reference it never
// allocates, never synchronises the device and returns a status code instead of
// throwing.
extern "C" int caus_gpu_ehrlich_aberth(void* stream, void** buffers, const char* opaque,
                                       size_t opaque_len) {
    if (opaque_len != sizeof(CausEhrlichAberthDescriptor) || !buffers) return CAUS_ERR_ARG;
    CausEhrlichAberthDescriptor d;
    __builtin_memcpy(&d, opaque, sizeof(d));
    return caus_roots(stream, (const double*)buffers[0], nullptr, (double*)buffers[1], nullptr,
                      nullptr, d.size, d.deg, d.itmax,
                      d.compensated ? CAUS_MODE_FP64_COMP : CAUS_MODE_FP64, CAUS_INIT_REFERENCE);
}

// The reference target's exact signature, void return (kernels.h:19-20): an XLA
// custom-call target cannot return a value, so the status of the last call on this
// host thread is kept for caus_gpu_ehrlich_aberth_status.
static thread_local int g_gpu_ea_status = CAUS_OK;

extern "C" void caus_gpu_ehrlich_aberth_void(void* stream, void** buffers, const char* opaque,
                                             size_t opaque_len) {
    g_gpu_ea_status = caus_gpu_ehrlich_aberth(stream, buffers, opaque, opaque_len);
}

extern "C" int caus_gpu_ehrlich_aberth_status(void) {
    const int s = g_gpu_ea_status;
    g_gpu_ea_status = CAUS_OK;
    return s;
}

namespace caus {
#define CAUS_DECL_VJP(D) \
    extern template int roots_vjp_for_degree<D>(hipStream_t, const cd*, const cd*, const cd*, cd*, int64_t);
CAUS_DECL_VJP(1) CAUS_DECL_VJP(2) CAUS_DECL_VJP(3) CAUS_DECL_VJP(4) CAUS_DECL_VJP(5) CAUS_DECL_VJP(6)
CAUS_DECL_VJP(7) CAUS_DECL_VJP(8) CAUS_DECL_VJP(9) CAUS_DECL_VJP(10) CAUS_DECL_VJP(11) CAUS_DECL_VJP(12)
#undef CAUS_DECL_VJP
}  // namespace caus

// Coefficient gradient of the root map (the VJP of ehrlich_aberth_primitive.py:254-324).
extern "C" int caus_roots_vjp(void* stream, const double* coeffs, const double* roots,
                              const double* g_roots, double* g_coeffs, int64_t N, int32_t deg) {
    if (N < 0 || (N > 0 && (!coeffs || !roots || !g_roots || !g_coeffs))) return CAUS_ERR_ARG;
    if (N == 0) return CAUS_OK;
    hipStream_t s = (hipStream_t)stream;
    const cd* c = reinterpret_cast<const cd*>(coeffs);
    const cd* r = reinterpret_cast<const cd*>(roots);
    const cd* g = reinterpret_cast<const cd*>(g_roots);
    cd* gc = reinterpret_cast<cd*>(g_coeffs);
    switch (deg) {
#define CAUS_DEG(D) \
    case D:         \
        return roots_vjp_for_degree<D>(s, c, r, g, gc, N);
        CAUS_DEG(1)
        CAUS_DEG(2)
        CAUS_DEG(3)
        CAUS_DEG(4)
        CAUS_DEG(5)
        CAUS_DEG(6)
        CAUS_DEG(7)
        CAUS_DEG(8)
        CAUS_DEG(9)
        CAUS_DEG(10)
        CAUS_DEG(11)
        CAUS_DEG(12)
#undef CAUS_DEG
        default:
            return CAUS_ERR_ARG;
    }
}

// caus_roots_vjp as an XLA custom-call target beside caus_gpu_ehrlich_aberth_void:
// buffers = {coeffs, roots, g_roots, g_coeffs}, opaque = the descriptor (size, deg used).
extern "C" void caus_gpu_ehrlich_aberth_vjp_void(void* stream, void** buffers,
                                                 const char* opaque, size_t opaque_len) {
    if (opaque_len != sizeof(CausEhrlichAberthDescriptor) || !buffers) {
        g_gpu_ea_status = CAUS_ERR_ARG;
        return;
    }
    CausEhrlichAberthDescriptor d;
    __builtin_memcpy(&d, opaque, sizeof(d));
    g_gpu_ea_status = caus_roots_vjp(stream, (const double*)buffers[0], (const double*)buffers[1],
                                     (const double*)buffers[2], (double*)buffers[3], d.size, d.deg);
}

extern "C" int caus_abi_version(void) { return CAUS_ABI_VERSION; }
