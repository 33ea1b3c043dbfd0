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

/* The same with the reference target's exact signature and void return
 * (lib/ehrlich_aberth/kernels.h:19-20, registered by gpu_ops.cc:14-26): what an XLA
 * custom-call registration binds.  The status code of the last call on the calling host
 * thread is read (and reset to CAUS_OK) by caus_gpu_ehrlich_aberth_status. */
void caus_gpu_ehrlich_aberth_void(void* stream, void** buffers, const char* opaque,
                                  size_t opaque_len);
int caus_gpu_ehrlich_aberth_status(void);

/* Coefficient gradient of the root map: the adjoint of the implicit-function JVP
 * _ehrlich_aberth_jvp (src/caustics/ehrlich_aberth_primitive.py:254-324),
 * dz_j/dp_k = -z_j^k / p'(z_j).  With cotangents g_roots (N, deg) it writes
 *   g_coeffs[i][k] = sum_j g_roots[i][j] * conj(dz_j/dp_k)     ((N, deg+1), ascending),
 * the conjugate-Wirtinger convention of torch.autograd.  roots are the converged roots
 * returned by caus_roots for the same coefficients. */
int caus_roots_vjp(void* stream, const double* coeffs, const double* roots, const double* g_roots,
                   double* g_coeffs, int64_t N, int32_t deg);

/* caus_roots_vjp as a custom-call target of the same shape as
 * caus_gpu_ehrlich_aberth_void, for the transpose of the primitive's JVP
 * (ehrlich_aberth_primitive.py:254-324) on the JAX side: buffers = {coeffs, roots,
 * g_roots, g_coeffs}, opaque = CausEhrlichAberthDescriptor bytes (size and deg are
 * used).  Status as caus_gpu_ehrlich_aberth_status. */
void caus_gpu_ehrlich_aberth_vjp_void(void* stream, void** buffers, const char* opaque,
                                      size_t opaque_len);

/* Point-source magnification, fused (reference: mag_point_source,
 * src/caustics/point_source.py:1762-1830, with _images_point_source :1652-1709,
 * lens_eq_det_jac :1558-1580 and the coefficients :17-86 / :89-1479).
 *   w        (N,) complex source positions (frame of the reference API)
 *   params   rows of (s, q) for nlenses=2, (s, q, q3, r3, psi) for nlenses=3, unused for 1;
 *            row i at params + i*params_stride (params_stride = 0 broadcasts one row)
 *   mag      (N,) output;  status (N,) uint8 or NULL (1 = every root converged)
 * mode CAUS_MODE_*; initial estimates are always CAUS_INIT_COMPLEX (order-free sum). */
int caus_mag_point_source(void* stream, int32_t nlenses, const double* w, int64_t N,
                          const double* params, int64_t params_stride, double* mag,
                          uint8_t* status, int32_t itmax, int32_t mode);

/* Forward + VJP in one pass (one root solve per evaluation).  For cotangents g_mag (N,):
 *   g_w      (N,) complex or NULL:  g * (dmag/dRe w + i dmag/dIm w)
 *   g_params (N, P) or NULL:        g * dmag/dparam_k per evaluation (P = 2 or 5)
 * The derivative is the implicit one of the lens equation at each real image
 * (equal to the reference's polynomial-root JVP, ehrlich_aberth_primitive.py:254-324,
 * chained through the coefficients); the image filter is piecewise constant. */
int caus_mag_point_source_vjp(void* stream, int32_t nlenses, const double* w, int64_t N,
                              const double* params, int64_t params_stride, const double* g_mag,
                              double* mag, double* g_w, double* g_params, uint8_t* status,
                              int32_t itmax, int32_t mode);

/* Point-source light curves along rectilinear trajectories w(t) = e^{-i alpha}(u0 + i (t-t0)/tE)
 * (src/caustics/trajectory.py:152-155 without parallax), L light curves x T epochs.
 *   t          (T,) epochs;  lc_params (L, 4+P) rows (t0, u0, tE, alpha, lens params...)
 *   mag        (L, T) output;  status (L, T) uint8 or NULL
 *   g_mag      (L, T) cotangents or NULL (forward only)
 *   g_params   (L, 4+P) output: sum over epochs of g_mag * dmag/dparam, reduced in a
 *              fixed order (deterministic); needs workspace of
 *              caus_lightcurve_workspace_bytes(nlenses, T, L) bytes.  L <= 65535. */
int64_t caus_lightcurve_workspace_bytes(int32_t nlenses, int64_t T, int64_t L);
int caus_lightcurve_point_source(void* stream, int32_t nlenses, const double* t, int64_t T,
                                 const double* lc_params, int64_t L, const double* g_mag,
                                 double* mag, double* g_params, double* workspace,
                                 int64_t workspace_bytes, uint8_t* status, int32_t itmax,
                                 int32_t mode);

/* caus_lightcurve_point_source without a workspace argument: it allocates nothing and
 * takes no caller memory beyond inputs and outputs.  No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The per-chunk partial gradient sums
 * live in a static device array of the library, CAUS_LC_NOALLOC_MAX_BYTES (8 MiB) per
 * GPU, held until the process ends: batches whose caus_lightcurve_workspace_bytes exceed
 * it are refused with CAUS_ERR_ARG before anything is enqueued.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (concurrent streams, threads) race;
 * calls ordered on one stream do not.  Forward-only calls (g_mag NULL) use no scratch. */
#define CAUS_LC_NOALLOC_MAX_BYTES (8 << 20)
int caus_lightcurve_point_source_noalloc(void* stream, int32_t nlenses, const double* t, int64_t T,
                                         const double* lc_params, int64_t L, const double* g_mag,
                                         double* mag, double* g_params, uint8_t* status,
                                         int32_t itmax, int32_t mode);

/* Hexadecapole magnification and contour gate of mag() (src/caustics/lightcurve.py:168-227):
 * _mag_hexadecapole (src/caustics/multipole.py:208-244) with linear limb darkening u1
 * (mag() itself passes u1 = 0, lightcurve.py:215), _caustics_proximity_test
 * (lightcurve.py:23-75) and, for q < 0.01, _planetary_caustic_test (lightcurve.py:78-85).
 *   rho      source radius per point at rho + i*rho_stride
 *   mu_hex   (P,) hexadecapole magnification; delta_mu (P,) or NULL: sum |dq|+|dh|
 *   use_hex  (P,) uint8: 1 where the hexadecapole value is accepted, 0 where the full
 *            contour integral is needed (always 0 for nlenses=3; |w| > 2 rho for one lens) */
int caus_hexadecapole_gate(void* stream, int32_t nlenses, const double* w, int64_t P,
                           const double* params, int64_t params_stride, const double* rho,
                           int64_t rho_stride, double u1, double* mu_hex, double* delta_mu,
                           uint8_t* use_hex, uint8_t* status, int32_t itmax);

/* Reverse mode of the hexadecapole magnification (the derivative the reference gets by
 * JAX autodiff of _mag_hexadecapole, src/caustics/multipole.py:208-244, on the points
 * mag() accepts, src/caustics/lightcurve.py:215-254).  Arguments as for
 * caus_hexadecapole_gate, plus
 *   g_mu   (P,) cotangent of mu_hex
 *   mu_hex (P,) out: the hexadecapole magnification (same value as the gate's)
 *   g_out  (P, 4 + n_params) out: g_mu * d mu_hex / d(Re w, Im w, rho, u1, lens params),
 *          n_params = 0, 2, 5 for nlenses = 1, 2, 3 */
int caus_hexadecapole_vjp(void* stream, int32_t nlenses, const double* w, int64_t P,
                          const double* params, int64_t params_stride, const double* rho,
                          int64_t rho_stride, double u1, const double* g_mu, double* mu_hex,
                          double* g_out, uint8_t* status, int32_t itmax);

/* Extended-source magnification by contour integration for a batch of sources, as one
 * fixed sequence of launches over the batch (reference: mag_extended_source, src/caustics/extended_source.py:730-901, with
 * _images_of_source_limb :56-153, segments and stitching :156-727 and the integrals of
 * src/caustics/integrate.py:18-121).
 *   w0 (S,) complex source centres; rho (S,); params rows as for caus_mag_point_source
 *   npts_limb (S,) int32 per-source limb resolution (the reference's npts_limb) or NULL to
 *            use npts_max for every source; each value must lie in [4, npts_max]
 *            (out-of-range values are clamped and flagged in status), or be 0: the source
 *            is skipped (no work; mag 0, gradient 0, status 1), so a caller can leave
 *            sources out on the device without reordering the batch
 *   npts_max  largest npts_limb, <= 4096 (sets the workspace size)
 *   limb_darkening 0: uniform source (trapezoid Green integral); 1: linear limb darkening
 *            with coefficient u1[i*u1_stride] and npts_ld Gauss-Legendre points per line
 *            integral (P, Q of Dominik 1998), npts_ld in [2, 256]
 *   mag (S,) output; status (S,) uint8 or NULL (0: a root solve did not converge or
 *            npts_limb was out of range)
 *   workspace: caus_extended_workspace_bytes(nlenses, S, npts_max) bytes of device memory
 * mode CAUS_MODE_FP64, CAUS_MODE_FP64_COMP (the limb solves run the compensated stage,
 * the reference's roots_compensated=True, extended_source.py:750,839) or CAUS_MODE_CF32.
 * Every limb solve ends with a compensated polish to the doubles nearest the exact
 * roots, so the three modes see the same images. */
int64_t caus_extended_workspace_bytes(int32_t nlenses, int64_t S, int32_t npts_max);
int caus_mag_extended_source(void* stream, int32_t nlenses, const double* w0, int64_t S,
                             const double* rho, const double* params, int64_t params_stride,
                             const int32_t* npts_limb, int32_t npts_max, int32_t limb_darkening,
                             const double* u1, int64_t u1_stride, int32_t npts_ld, double* mag,
                             uint8_t* status, void* workspace, int64_t workspace_bytes,
                             int32_t itmax, int32_t mode);

/* Launch schedule of the contour pipeline's grid-stride kernels (limb solves, added
 * points): at most n workgroups of 256 lanes per CU (default 8, 1 <= n <= 64).  No
 * reference counterpart (the reference maps over sources sequentially,
 * lightcurve.py:245-254); set per workload by caustics_amd.autotune from its tuned
 * table.  Process-wide; read when a batch is launched (a captured graph keeps the
 * grid it was captured with).  Results do not depend on it. */
int caus_ext_set_flat_blocks_per_cu(int32_t n);

/* Forward + VJP of caus_mag_extended_source in one pass.  For cotangents g_mag (S,) it
 * writes g_params (S, 4+P): g * dmag/d(Re w0, Im w0, rho, u1, lens params...), the
 * derivative of the discretised integral as the reference's autodiff takes it (limb
 * angles, image rows and segment structure fixed; every contour point differentiated
 * implicitly through the lens equation; Gauss-Legendre integrals with their moving end
 * and split points). */
int caus_mag_extended_source_vjp(void* stream, int32_t nlenses, const double* w0, int64_t S,
                                 const double* rho, const double* params, int64_t params_stride,
                                 const int32_t* npts_limb, int32_t npts_max,
                                 int32_t limb_darkening, const double* u1, int64_t u1_stride,
                                 int32_t npts_ld, const double* g_mag, double* mag,
                                 double* g_params, uint8_t* status, void* workspace,
                                 int64_t workspace_bytes, int32_t itmax, int32_t mode);

/* caus_mag_extended_source_vjp (g_mag may be NULL for the forward only) that also adds
 * work counts to work[5] (uint64, caller-zeroed): FP64 root-iterations, FP32
 * root-iterations, compensated root-iterations, limb-point solves, contour points.  The
 * counters cost atomics; they exist to price a run against the VALU roofline.  No
 * counterpart in the reference. */
int caus_mag_extended_source_work(void* stream, int32_t nlenses, const double* w0, int64_t S,
                                  const double* rho, const double* params,
                                  int64_t params_stride, const int32_t* npts_limb,
                                  int32_t npts_max, int32_t limb_darkening, const double* u1,
                                  int64_t u1_stride, int32_t npts_ld, const double* g_mag,
                                  double* mag, double* g_params, uint8_t* status,
                                  void* workspace, int64_t workspace_bytes, int32_t itmax,
                                  int32_t mode, uint64_t* work);

/* Greedy continuation of root sets along a sequence: for each of nchains chains of npts
 * sets of n points (pts (nchains, npts, n) complex), out[k] = pts[k] permuted so that
 * each row continues the previous set by the reference's match_points rule
 * (src/caustics/utils.py:15-40, chained as in critical_and_caustic_curves,
 * src/caustics/point_source.py:1628-1640).  n in {2, 4, 5, 6, 10}; pts != out. */
int caus_match_chain(void* stream, const double* pts, double* out, int64_t npts, int32_t n,
                     int64_t nchains);

/* Images of point sources in the lens frame (reference: _images_point_source,
 * src/caustics/point_source.py:1652-1709) and their sequential, warm-started form
 * (_images_point_source_sequential, :1711-1759).
 *   w (B, N) complex source positions in the lens frame
 *   lens (B rows, stride lens_stride doubles): (a, e1) for nlenses 2, (a, e1, e2, Re r3,
 *     Im r3) for 3 (lens 1 at +a with mass e1, lens 2 at -a, lens 3 at r3); unused for 1
 *   z_init (B, N, n) complex or NULL: initial roots (the reference's custom_init); with
 *     sequential = 1 only point 0 of each sequence uses them
 *   sequential 0: every position solved on its own; 1: each of the B sequences solved in
 *     order, position k > 0 started from position k-1's roots, so that root j follows one
 *     image along the sequence
 *   z (B, N, n) complex out; mask (B, N, n) uint8 out, 1 where |lens_eq(z) - w| < 1e-6;
 *   status (B, N) uint8 or NULL (1 = converged).  n = 2, 5, 10 for nlenses 1, 2, 3. */
int caus_images_point_source(void* stream, int32_t nlenses, const double* w, int64_t B,
                             int64_t N, const double* lens, int64_t lens_stride,
                             const double* z_init, double* z, uint8_t* mask, uint8_t* status,
                             int32_t itmax, int32_t mode, int32_t sequential);

/* Flux-marginalised log-likelihood with a diagonal data covariance for B magnification
 * curves that share one set of observations (reference: marginalized_log_likelihood,
 * src/caustics/linalg.py:10-70, dense_covariance=False; the lists of the reference are
 * the nseg segments here).
 *   A (B, n) magnifications, row stride lda >= n; the n epochs are nseg observatory
 *     segments, epochs seg_off[k] .. seg_off[k+1]-1 (seg_off (nseg+1,) int64 on the
 *     device, seg_off[0] = 0, seg_off[nseg] = n, non-decreasing)
 *   fobs (n,) observed fluxes, c_inv (n,) the diagonal of the inverse covariance
 *   ll (B,) out: sum over segments of -chi2/2 + log det(2 pi Sigma)/2 (linalg.py:68)
 *   beta (B, nseg, 2) out or NULL: the least-squares (F_s, F_b) of each segment
 *   g_A (B, n) out or NULL, row stride ldg >= n: g_ll[b] * d ll_b / d A_bi (g_ll NULL:
 *     cotangent 1); the gradient the reference's autodiff gives with respect to A
 * A segment whose normal matrix is singular (fewer than two distinct magnifications)
 * gives a non-finite ll, as the reference's solve does.  B <= 2^26. */
int caus_marginalized_log_likelihood(void* stream, const double* A, int64_t B, int64_t n,
                                     int64_t lda, const double* fobs, const double* c_inv,
                                     const int64_t* seg_off, int32_t nseg, double* ll,
                                     double* beta, const double* g_ll, double* g_A, int64_t ldg);

/* Source trajectories with annual parallax, and their reverse mode.
 * Replaces AnnualParallaxTrajectory.compute and _compute_delta_sun_position_and_velocity
 * (src/caustics/trajectory.py:104-157), for L parameter sets at once.
 *   t (T,) epochs; xp (M,) increasing grid, M >= 2 (the reference's daily t_jpl, :36);
 *   sun_tab (4, M): rows s_e, s_n, sdot_e, sdot_n, the Sun's position and velocity
 *     relative to the Earth projected east / north on the plane of the sky (:80-102)
 *   params (L, 5): t0, u0, tE, then pi_EE, pi_EN (parametrization 0, "cartesian") or
 *     pi_E, psi (parametrization 1, "polar") (:137-146)
 *   w (L, T) complex out or NULL: u_e + i u_n (:149-155), the positions interpolated as
 *     jax.numpy.interp does (clamped to the end values outside xp)
 *   g_w (L, T) complex cotangents (dL/dRe w + i dL/dIm w) or NULL; with g_w, g_params
 *     (L, 5) out receives dL/dparams in the order of params.  The cartesian gradient
 *     is undefined (non-finite) at pi_E = 0, where psi is.
 * L <= 2^31 - 1. */
int caus_trajectory_annual_parallax(void* stream, const double* t, int64_t T, const double* xp,
                                    const double* sun_tab, int64_t M, const double* params,
                                    int64_t L, int32_t parametrization, double* w,
                                    const double* g_w, double* g_params);

#ifdef __cplusplus
}
#endif

#endif /* CAUSTICS_HIP_H */
