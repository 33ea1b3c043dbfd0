// Root-solver kernel template and per-degree launcher, instantiated once per degree
// in roots_dNN.hip so the degrees compile in parallel.
#pragma once
#include "caustics_hip.h"
#include "cf32.hpp"
#include "solver.hpp"

namespace caus {

constexpr int kRootsBlock = 256;

template <int DEG, int MODE, bool REFCOMPAT>
__global__ __launch_bounds__(kRootsBlock) void k_roots(const cd* __restrict__ coeffs,
                                                       const cd* __restrict__ roots_init,
                                                       cd* __restrict__ roots,
                                                       int32_t* __restrict__ stats,
                                                       uint8_t* __restrict__ status, int64_t N,
                                                       int itmax) {
    for (int64_t i = (int64_t)blockIdx.x * kRootsBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kRootsBlock) {
    cd p[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) p[k] = coeffs[i * (DEG + 1) + k];
    prescale_coeffs<DEG>(p);
    cd z[DEG];
    const bool custom = roots_init != nullptr;
    if (custom) {
#pragma unroll
        for (int k = 0; k < DEG; ++k) z[k] = roots_init[i * DEG + k];
    }
    SolveStats st;
    bool ok;
    if (MODE == MODE_CF32)
        ok = solve_cf32<DEG, true>(p, z, itmax, custom, st);
    else
        ok = solve_fp64<DEG, MODE, REFCOMPAT, true>(p, z, itmax, custom, st);
#pragma unroll
    for (int k = 0; k < DEG; ++k) roots[i * DEG + k] = z[k];
    if (status) status[i] = ok ? 1 : 0;
    if (stats) {
        stats[4 * i + 0] = st.sweeps;
        stats[4 * i + 1] = st.it64;
        stats[4 * i + 2] = st.itcomp;
        stats[4 * i + 3] = st.it32;
    }
    }
}

// Initial estimates for every polynomial (init_est.h:57-102), written to `roots`.
template <int DEG, bool REFCOMPAT>
__global__ __launch_bounds__(kRootsBlock) void k_roots_init(const cd* __restrict__ coeffs,
                                                            cd* __restrict__ roots, int64_t N) {
    for (int64_t i = (int64_t)blockIdx.x * kRootsBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kRootsBlock) {
    cd p[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) p[k] = coeffs[i * (DEG + 1) + k];
    prescale_coeffs<DEG>(p);
    double al[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) al[k] = cabs_safe(p[k]);
    cd z[DEG];
    init_est<DEG, REFCOMPAT>(al, z);
#pragma unroll
    for (int k = 0; k < DEG; ++k) roots[i * DEG + k] = z[k];
    }
}

// Persistent solve with lane refill: each lane walks the polynomials i, i + L, i + 2L,
// ... (L = lanes in the grid) and starts its next polynomial as soon as the current one
// has converged, so a wave no longer waits for its slowest polynomial at every step.
// Initial estimates come from `init` (k_roots_init output or the caller's roots_init).
template <int DEG, int MODE, bool REFCOMPAT>
__global__ __launch_bounds__(kRootsBlock) void k_roots_persistent(
    const cd* __restrict__ coeffs, const cd* init, cd* roots, int32_t* __restrict__ stats,
    uint8_t* __restrict__ status, int64_t N, int itmax) {
    const int64_t L = (int64_t)gridDim.x * kRootsBlock;
    int64_t i = (int64_t)blockIdx.x * kRootsBlock + threadIdx.x;
    cd p[DEG + 1], z[DEG];
    EAState<DEG, MODE, REFCOMPAT, true> S;
    SolveStats st{0, 0, 0, 0};
    bool active = i < N;
    if (active) {
#pragma unroll
        for (int k = 0; k <= DEG; ++k) p[k] = coeffs[i * (DEG + 1) + k];
        prescale_coeffs<DEG>(p);
#pragma unroll
        for (int k = 0; k < DEG; ++k) z[k] = init[i * DEG + k];
        S.start(p, z, true);
    }
    while (__ballot(active) != 0ull) {
        if (active) {
            const bool done = S.sweep(p, z, st);
            if (done || st.sweeps >= itmax) {
#pragma unroll
                for (int k = 0; k < DEG; ++k) roots[i * DEG + k] = z[k];
                if (status) status[i] = (done && roots_finite<DEG>(z)) ? 1 : 0;
                if (stats) {
                    stats[4 * i + 0] = st.sweeps;
                    stats[4 * i + 1] = st.it64;
                    stats[4 * i + 2] = st.itcomp;
                    stats[4 * i + 3] = 0;
                }
                i += L;
                active = i < N;
                if (active) {
#pragma unroll
                    for (int k = 0; k <= DEG; ++k) p[k] = coeffs[i * (DEG + 1) + k];
                    prescale_coeffs<DEG>(p);
#pragma unroll
                    for (int k = 0; k < DEG; ++k) z[k] = init[i * DEG + k];
                    S.start(p, z, true);
                    st = SolveStats{0, 0, 0, 0};
                }
            }
        }
    }
}

// Fast FP64 form with the converged roots moved behind the unconverged ones at the end
// of every sweep in which one converged (a stable partition, so the unconverged roots
// keep their order and a sweep still visits them in the Gauss-Seidel order of
// EAState::sweep).  The sweep loop then runs root slot j only while some lane has more
// than j unconverged roots, instead of whenever any lane still needs its own root j.
// The partition goes through LDS (one column per lane, written at the lane's target
// slots and read back in slot order); `perm` maps slots to the original root indices
// for the output.  The Aberth sums add the same terms in slot order, so the roots agree
// with the uncompacted sweep to rounding.
template <int DEG>
struct CompactEA {
    double al[DEG + 1];  // alpha' weights of the stopping test
    unsigned cx;         // converged slots; every slot >= na is converged
    int na;              // unconverged slots at the start of the sweep
    uint64_t perm;       // slot -> original root index, 4 bits per slot

    // a[i] = |p_i|; z gets the initial estimates of init_est unless given
    __device__ __forceinline__ void start(const double (&a)[DEG + 1], cd (&z)[DEG], bool given) {
        if (!given) init_est<DEG, false>(a, z);
#pragma unroll
        for (int i = 0; i <= DEG; ++i) al[i] = a[i] * fma(ALPHA_W, (double)i, 1.0);
        cx = 0;
        na = DEG;
        perm = 0;
#pragma unroll
        for (int i = 0; i < DEG; ++i) perm |= (uint64_t)i << (4 * i);
    }
    // zs: this lane's LDS column (slot k at zs[k * kRootsBlock])
    __device__ __forceinline__ bool sweep(const cd (&p)[DEG + 1], cd (&z)[DEG], SolveStats& st,
                                          double2* zs) {
        constexpr unsigned ALL = (1u << DEG) - 1u;
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            if (j < na && !((cx >> j) & 1u)) {
                ++st.it64;
                if (fast_step<DEG, true>(p, al, z, j)) cx |= 1u << j;
            }
        }
        ++st.sweeps;
        if (cx == ALL) return true;
        if (cx & ((1u << na) - 1u)) {
            const unsigned act = ~cx & ALL;
            const int nact = __builtin_popcount(act);
            uint64_t np = 0;
#pragma unroll
            for (int j = 0; j < DEG; ++j) {
                const unsigned below = (1u << j) - 1u;
                const int d = ((act >> j) & 1u) ? __builtin_popcount(act & below)
                                                : nact + __builtin_popcount(cx & below);
                zs[d * kRootsBlock] = make_double2(z[j].re, z[j].im);
                np |= ((perm >> (4 * j)) & 0xFull) << (4 * d);
            }
#pragma unroll
            for (int k = 0; k < DEG; ++k) {
                const double2 v = zs[k * kRootsBlock];
                z[k] = cd{v.x, v.y};
            }
            perm = np;
            na = nact;
            cx = ALL & ~((1u << nact) - 1u);
        }
        return false;
    }
};

// k_roots_persistent for the Fast FP64 form with CompactEA sweeps.  Launched with one
// polynomial per lane; lanes refill (stride L) only if the grid had to be capped.  With
// init null every lane computes its own initial estimates (init_est.h:57-102, the same
// function and inputs as k_roots_init, so the same doubles): all lanes start together,
// so this adds no divergence, and no estimates pass through HBM.  The estimates use the
// compile-time-indexed hull (init_est_static), so nothing goes through scratch; the
// kernel is held to 3 waves per SIMD (degree 10: 167 VGPRs, 3 dwords spilled, against
// 183 VGPRs and 2 waves unconstrained), measured 3% faster at degree 10 with results
// bit for bit the same.  A called, non-inlined init_est was more than twice as slow
// (its stack frame goes through scratch).
template <int DEG>
__global__ __launch_bounds__(kRootsBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_roots_compact(
    const cd* __restrict__ coeffs, const cd* init, cd* roots, int32_t* __restrict__ stats,
    uint8_t* __restrict__ status, int64_t N, int itmax) {
    __shared__ double2 zs[DEG * kRootsBlock];
    const int64_t L = (int64_t)gridDim.x * kRootsBlock;
    int64_t i = (int64_t)blockIdx.x * kRootsBlock + threadIdx.x;
    double2* col = zs + threadIdx.x;
    cd p[DEG + 1], z[DEG];
    CompactEA<DEG> S;
    SolveStats st{0, 0, 0, 0};
    bool active = i < N;
    auto load = [&](int64_t q) {
        double a[DEG + 1];
#pragma unroll
        for (int k = 0; k <= DEG; ++k) p[k] = coeffs[q * (DEG + 1) + k];
        prescale_coeffs<DEG>(p);
#pragma unroll
        for (int k = 0; k <= DEG; ++k) a[k] = cabs_safe(p[k]);
        if (init) {
#pragma unroll
            for (int k = 0; k < DEG; ++k) z[k] = init[q * DEG + k];
        }
        S.start(a, z, init != nullptr);
    };
    if (active) load(i);
    while (__ballot(active) != 0ull) {
        if (active) {
            const bool done = S.sweep(p, z, st, col);
            if (done || st.sweeps >= itmax) {
#pragma unroll
                for (int k = 0; k < DEG; ++k) roots[i * DEG + (int)((S.perm >> (4 * k)) & 0xF)] = z[k];
                if (status) status[i] = (done && roots_finite<DEG>(z)) ? 1 : 0;
                if (stats) {
                    stats[4 * i + 0] = st.sweeps;
                    stats[4 * i + 1] = st.it64;
                    stats[4 * i + 2] = 0;
                    stats[4 * i + 3] = 0;
                }
                i += L;
                active = i < N;
                if (active) {
                    load(i);
                    st = SolveStats{0, 0, 0, 0};
                }
            }
        }
    }
}

// degrees from which the compacted sweep pays for its partition
constexpr int kRootsCompactMinDeg = 5;

int roots_grid_limit();

// HIP limits a launch to 2^32 - 1 work-items; every root kernel strides over the batch
// (grid-stride or lane refill), so its grid is capped there.
constexpr int64_t kRootsMaxBlocks = 0xFFFFFFFFll / kRootsBlock;

template <int DEG, int MODE, bool RC>
int launch_roots(hipStream_t s, const cd* c, const cd* ri, cd* r, int32_t* st, uint8_t* ok,
                 int64_t N, int itmax) {
    const int64_t nb = (N + kRootsBlock - 1) / kRootsBlock;
    const int64_t blocks = nb < kRootsMaxBlocks ? nb : kRootsMaxBlocks;
    if (MODE == MODE_CF32) {
        hipLaunchKernelGGL((k_roots<DEG, MODE, RC>), dim3((unsigned)blocks), dim3(kRootsBlock), 0,
                           s, c, ri, r, st, ok, N, itmax);
        return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
    }
    if (itmax == 0) {
        // no sweep at all: the initial estimates are the result
        hipLaunchKernelGGL((k_roots<DEG, MODE, RC>), dim3((unsigned)blocks), dim3(kRootsBlock), 0,
                           s, c, ri, r, st, ok, N, itmax);
        return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
    }
    if (MODE == MODE_FP64 && !RC && DEG >= kRootsCompactMinDeg) {
        // one polynomial per lane: a partition helps while the lanes of a wave are at
        // similar stages, which refilled lanes are not; the workgroup dispatcher then
        // balances the load (measured: 4% faster at degree 5 than with refill).  The
        // kernel computes the initial estimates itself when none are given.
        hipLaunchKernelGGL((k_roots_compact<DEG>), dim3((unsigned)blocks), dim3(kRootsBlock), 0, s, c,
                           ri, r, st, ok, N, itmax);
        return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
    }
    const cd* init = ri;
    if (!ri) {
        hipLaunchKernelGGL((k_roots_init<DEG, RC>), dim3((unsigned)blocks), dim3(kRootsBlock), 0, s,
                           c, r, N);
        init = r;
    }
    const int64_t pb = blocks < roots_grid_limit() ? blocks : roots_grid_limit();
    hipLaunchKernelGGL((k_roots_persistent<DEG, MODE, RC>), dim3((unsigned)pb), dim3(kRootsBlock),
                       0, s, c, init, r, st, ok, N, itmax);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}


// Adjoint of the root map (implicit function theorem, ehrlich_aberth_primitive.py:
// 305-319): dz_j/dp_k = -z_j^k / p'(z_j).  For cotangents g_j the coefficient gradient
// in the conjugate-Wirtinger convention is  gc_k = sum_j g_j conj(dz_j/dp_k).
template <int DEG>
__global__ __launch_bounds__(kRootsBlock) void k_roots_vjp(const cd* __restrict__ coeffs,
                                                           const cd* __restrict__ roots,
                                                           const cd* __restrict__ g,
                                                           cd* __restrict__ gc, int64_t N) {
    for (int64_t i = (int64_t)blockIdx.x * kRootsBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kRootsBlock) {
    cd p[DEG + 1], acc[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) {
        p[k] = coeffs[i * (DEG + 1) + k];
        acc[k] = cd{0.0, 0.0};
    }
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
        const cd zj = roots[i * DEG + j];
        const cd gj = g[i * DEG + j];
        cd d = (double)DEG * p[DEG];
#pragma unroll
        for (int k = DEG - 1; k >= 1; --k) d = cfma(d, zj, (double)k * p[k]);
        const cd u = -(gj * conj(crcp<2>(d)));
        const cd zc = conj(zj);
        cd pw = cd{1.0, 0.0};
#pragma unroll
        for (int k = 0; k <= DEG; ++k) {
            acc[k] = cfma(u, pw, acc[k]);
            pw = pw * zc;
        }
    }
#pragma unroll
    for (int k = 0; k <= DEG; ++k) gc[i * (DEG + 1) + k] = acc[k];
    }
}

template <int DEG>
int roots_vjp_for_degree(hipStream_t s, const cd* c, const cd* r, const cd* g, cd* gc, int64_t N) {
    const int64_t nb = (N + kRootsBlock - 1) / kRootsBlock;
    const int64_t blocks = nb < kRootsMaxBlocks ? nb : kRootsMaxBlocks;
    hipLaunchKernelGGL((k_roots_vjp<DEG>), dim3((unsigned)blocks), dim3(kRootsBlock), 0, s, c, r, g,
                       gc, N);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}

// Launch for one degree; defined in roots_dNN.hip.
template <int DEG>
int roots_for_degree(hipStream_t s, const cd* c, const cd* ri, cd* r, int32_t* st, uint8_t* ok,
                     int64_t N, int itmax, int mode, int init_mode);

#define CAUS_DEFINE_ROOTS_DEGREE(DEG)                                                           \
    template <>                                                                                 \
    int roots_for_degree<DEG>(hipStream_t s, const cd* c, const cd* ri, cd* r, int32_t* st,    \
                              uint8_t* ok, int64_t N, int itmax, int mode, int init_mode) {     \
        const bool rc = init_mode == CAUS_INIT_REFERENCE;                                       \
        switch (mode) {                                                                         \
            case CAUS_MODE_FP64:                                                                \
                return rc ? launch_roots<DEG, MODE_FP64, true>(s, c, ri, r, st, ok, N, itmax)   \
                          : launch_roots<DEG, MODE_FP64, false>(s, c, ri, r, st, ok, N, itmax); \
            case CAUS_MODE_FP64_COMP:                                                           \
                return rc ? launch_roots<DEG, MODE_FP64_COMP, true>(s, c, ri, r, st, ok, N,     \
                                                                    itmax)                      \
                          : launch_roots<DEG, MODE_FP64_COMP, false>(s, c, ri, r, st, ok, N,    \
                                                                     itmax);                    \
            case CAUS_MODE_CF32:                                                                \
                return launch_roots<DEG, MODE_CF32, false>(s, c, ri, r, st, ok, N, itmax);      \
            default:                                                                            \
                return CAUS_ERR_ARG;                                                            \
        }                                                                                       \
    }                                                                                           \
    template int roots_vjp_for_degree<DEG>(hipStream_t, const cd*, const cd*, const cd*, cd*,   \
                                           int64_t);

}  // namespace caus
