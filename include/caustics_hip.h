/*
 * caustics_hip.h — C ABI of the MI355X (gfx950) microlensing magnification hot path.
 *
 * Every entry point:
 *   - takes a HIP stream (hipStream_t passed as void*, NULL = default stream) and
 *     device pointers that the caller owns; complex numbers are interleaved
 *     (re, im) doubles, arrays are C-contiguous;
 *   - only enqueues kernels: it never allocates, never synchronises the device and
 *     never prints; results are ready when the stream reaches that point;
 *   - returns CAUS_OK (0) when the launch was enqueued, CAUS_ERR_ARG (-1) for an
 *     invalid argument (nothing enqueued) and CAUS_ERR_LAUNCH (-2) when the HIP
 *     runtime rejected the launch.  Per-item convergence is reported in `status`
 *     arrays (1 = converged), never by aborting.
 */
#ifndef CAUSTICS_HIP_H
#define CAUSTICS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAUS_ABI_VERSION 1

#define CAUS_OK 0
#define CAUS_ERR_ARG (-1)
#define CAUS_ERR_LAUNCH (-2)

/* Precision forms of the root solver. */
#define CAUS_MODE_FP64 0      /* plain Ehrlich-Aberth, ehrlich_aberth.h:60-150        */
#define CAUS_MODE_FP64_COMP 1 /* + compensated stage, ehrlich_aberth.h:153-292         */
#define CAUS_MODE_CF32 2      /* FP32 iterations + compensated FP32 sweep + FP64 polish */

/* Initial estimates. */
#define CAUS_INIT_REFERENCE 0 /* real-axis estimates of init_est.h:95 as compiled, reversed
                                 branch for |z|>1: reproduces the reference root order   */
#define CAUS_INIT_COMPLEX 1   /* r(cos t + i sin t) estimates, forward polynomial only    */

/* Descriptor of the reference's GPU custom call, lib/ehrlich_aberth/kernels.h:11-17. */
typedef struct {
    int size;
    int deg;
    int itmax;
#ifdef __cplusplus
    bool compensated;
#else
    _Bool compensated;
#endif
} CausEhrlichAberthDescriptor;

int caus_abi_version(void);

/* Batched all-roots solve.
 * Replaces cpu_ehrlich_aberth (lib/ehrlich_aberth/cpu_ops.cc:15-81) and
 * gpu_ehrlich_aberth (lib/ehrlich_aberth/kernels.cc.cu:52-104, kernels.h:19-20), i.e. the
 * JAX primitive ehrlich_aberth (src/caustics/ehrlich_aberth_primitive.py:98-106).
 *   coeffs      (N, deg+1) complex, ascending (coeffs[i][k] multiplies z^k)
 *   roots_init  (N, deg) complex or NULL; non-NULL = custom_init
 *   roots       (N, deg) complex output
 *   stats       (N, 4) int32 or NULL: sweeps, FP64 root-iterations, compensated /
 *               double-float root-iterations, FP32 root-iterations
 *   status      (N,) uint8 or NULL: 1 when every root converged within itmax sweeps
 * deg in [1, 12]; mode CAUS_MODE_*; init_mode CAUS_INIT_* (ignored for CF32).       */
int caus_roots(void* stream, const double* coeffs, const double* roots_init, double* roots,
               int32_t* stats, uint8_t* status, int64_t N, int32_t deg, int32_t itmax,
               int32_t mode, int32_t init_mode);

/* Same operation with the reference GPU custom-call signature
 * (lib/ehrlich_aberth/kernels.h:19-20): buffers = {coeffs, roots}, opaque = the bytes of
 * CausEhrlichAberthDescriptor.  Uses CAUS_INIT_REFERENCE. */
int caus_gpu_ehrlich_aberth(void* stream, void** buffers, const char* opaque, size_t opaque_len);

/* Coefficient gradient of the root map: the adjoint of the implicit-function JVP
 * _ehrlich_aberth_jvp (src/caustics/ehrlich_aberth_primitive.py:254-324),
 * dz_j/dp_k = -z_j^k / p'(z_j).  With cotangents g_roots (N, deg) it writes
 *   g_coeffs[i][k] = sum_j g_roots[i][j] * conj(dz_j/dp_k)     ((N, deg+1), ascending),
 * the conjugate-Wirtinger convention of torch.autograd.  roots are the converged roots
 * returned by caus_roots for the same coefficients. */
int caus_roots_vjp(void* stream, const double* coeffs, const double* roots, const double* g_roots,
                   double* g_coeffs, int64_t N, int32_t deg);

#ifdef __cplusplus
}
#endif

#endif /* CAUSTICS_HIP_H */
