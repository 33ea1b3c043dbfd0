// Extended-source magnification by contour integration: one workgroup per source.
//
// Reference (fbartolic/caustics, src/caustics/):
//   limb sampling + 10 gap-bisection rounds .... extended_source.py:56-153
//   image continuation (greedy match_points) ... extended_source.py:34-53, utils.py:15-40
//   segments: closure test, split, ordering .... extended_source.py:156-309
//   stitching of open segments ................. extended_source.py:312-667
//   contours, magnification .................... extended_source.py:670-901
//   trapezoid / limb-darkened P,Q integrals .... integrate.py:18-121, utils.py:89-100
//
// Layout per source (workspace): limb angles, images (point-major: z[k*DEG + row]),
// per-point flag words (bit r: real image in row r; bit 16+r: det J < 0), all double
// buffered for the insertions of the refinement rounds, plus one 64-bit word per point
// for the image permutations and a P/Q scratch for limb darkening.
//
// Identity of images along the limb: every limb point is solved from the default
// initial estimates in parallel (one lane per point); image rows are then chained with
// the reference's greedy nearest-point assignment.  Where the nearest-neighbour map
// between consecutive points is a bijection the greedy result does not depend on the
// row order, so those steps compose as permutations; the rare steps where two images
// compete are resolved with the greedy rule on the actual carried rows (wave 0).
// Segments are kept as lists of pieces (row, first point, length, direction) instead
// of dense zero-padded arrays; stitching reproduces the reference's ordering rules.
#pragma once
#include "caustics_hip.h"
#include "ps_impl.hpp"

namespace caus {

constexpr int kExtBlock = 256;    // block of the per-source integration kernel
constexpr int kExtWaveBlock = 64; // block of the per-source bookkeeping kernels (one wave)
constexpr int kExtMaxSeg = 32;     // >= 3 * (nlenses^2 + 1)
constexpr int kExtMaxPieces = 48;  // pieces of one stitched contour
constexpr int kExtSortMax = 4096;  // limb points per source (gaps sorted in LDS)

template <int NL>
constexpr int ext_degree() {
    return NL == 1 ? 2 : NL * NL + 1;
}

__host__ __device__ inline int ext_npts_init(int npts) { return npts / 2; }
__host__ __device__ inline int ext_n_add(int npts) { return (npts / 2) / 10; }
__host__ __device__ inline int ext_mmax(int npts) { return ext_npts_init(npts) + 10 * ext_n_add(npts); }

__host__ __device__ inline size_t ext_align(size_t x) { return (x + 255) & ~size_t(255); }

// bytes of workspace per source
__host__ __device__ inline size_t ext_ws_bytes(int deg, int mmax) {
    size_t b = 0;
    b += ext_align(2 * (size_t)mmax * sizeof(double));            // theta x2
    b += ext_align(2 * (size_t)mmax * deg * sizeof(cd));          // z x2
    b += ext_align(2 * (size_t)mmax * sizeof(uint32_t));          // flags x2
    b += ext_align((size_t)(mmax + 1) * sizeof(uint64_t));        // permutations / changepoints
    b += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double)); // P, Q
    b += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double)); // point cotangents
    b += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double)); // centroid cotangent parts
    b += ext_align(2 * (size_t)mmax * sizeof(int32_t));             // limb order x2
    b += ext_align(2 * (size_t)mmax * sizeof(uint64_t));            // gap keys x2
    return b;
}

struct ExtWS {
    double* theta[2];
    cd* z[2];
    uint32_t* flg[2];
    uint64_t* perm;
    double* pq;
    double* xy;
    double* yz;
    int32_t* ord[2];   // refinement: limb order -> index of the point in append order
    uint64_t* gk[2];   // refinement: key of the gap after each point in limb order
};

__device__ inline ExtWS ext_ws(char* base, int deg, int mmax) {
    ExtWS w;
    size_t o = 0;
    w.theta[0] = (double*)(base + o);
    w.theta[1] = w.theta[0] + mmax;
    o += ext_align(2 * (size_t)mmax * sizeof(double));
    w.z[0] = (cd*)(base + o);
    w.z[1] = w.z[0] + (size_t)mmax * deg;
    o += ext_align(2 * (size_t)mmax * deg * sizeof(cd));
    w.flg[0] = (uint32_t*)(base + o);
    w.flg[1] = w.flg[0] + mmax;
    o += ext_align(2 * (size_t)mmax * sizeof(uint32_t));
    w.perm = (uint64_t*)(base + o);
    o += ext_align((size_t)(mmax + 1) * sizeof(uint64_t));
    w.pq = (double*)(base + o);
    o += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double));
    w.xy = (double*)(base + o);
    o += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double));
    w.yz = (double*)(base + o);
    o += ext_align(2 * ((size_t)mmax * deg + 2) * sizeof(double));
    w.ord[0] = (int32_t*)(base + o);
    w.ord[1] = w.ord[0] + mmax;
    o += ext_align(2 * (size_t)mmax * sizeof(int32_t));
    w.gk[0] = (uint64_t*)(base + o);
    w.gk[1] = w.gk[0] + mmax;
    return w;
}

// ---- permutations packed as 4-bit entries ------------------------------------
__device__ __forceinline__ int pget(uint64_t p, int i) { return (int)((p >> (4 * i)) & 0xF); }
template <int DEG>
__device__ __forceinline__ uint64_t pident() {
    uint64_t p = 0;
#pragma unroll
    for (int i = 0; i < DEG; ++i) p |= (uint64_t)i << (4 * i);
    return p;
}
// (a o b)[i] = a[b[i]]
template <int DEG>
__device__ __forceinline__ uint64_t pcompose(uint64_t a, uint64_t b) {
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < DEG; ++i) r |= (uint64_t)pget(a, pget(b, i)) << (4 * i);
    return r;
}
constexpr uint64_t kConflictBit = 1ull << 63;

// greedy match_points(a, b): for i in order, nearest unused b (ties -> lowest index)
template <int DEG>
__device__ __forceinline__ uint64_t greedy_match(const cd (&a)[DEG], const cd (&b)[DEG]) {
    unsigned used = 0;
    uint64_t p = 0;
#pragma unroll
    for (int i = 0; i < DEG; ++i) {
        double best = 0.0;
        int bj = -1;
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            const double d = norm2(b[j] - a[i]);
            if (!((used >> j) & 1u) && (bj < 0 || d < best)) {
                best = d;
                bj = j;
            }
        }
        used |= 1u << bj;
        p |= (uint64_t)bj << (4 * i);
    }
    return p;
}

// nearest-neighbour map a_i -> argmin_j |b_j - a_i|; conflict bit when not a bijection
template <int DEG>
__device__ __forceinline__ uint64_t nn_map(const cd (&a)[DEG], const cd (&b)[DEG]) {
    unsigned hit = 0;
    bool conflict = false;
    uint64_t p = 0;
#pragma unroll
    for (int i = 0; i < DEG; ++i) {
        double best = norm2(b[0] - a[i]);
        int bj = 0;
#pragma unroll
        for (int j = 1; j < DEG; ++j) {
            const double d = norm2(b[j] - a[i]);
            if (d < best) {
                best = d;
                bj = j;
            }
        }
        conflict = conflict || ((hit >> bj) & 1u);
        hit |= 1u << bj;
        p |= (uint64_t)bj << (4 * i);
    }
    return conflict ? (p | kConflictBit) : p;
}

// ---- limb sampling ------------------------------------------------------------
// Initial limb angle k of m (extended_source.py:101-103): numpy/jax linspace(-pi, pi,
// m, endpoint=False) is start + k * step with the product and the sum rounded
// separately, then pi - 1e-8 as the last angle; no FMA contraction, so the angles are
// the same doubles.
__device__ __forceinline__ double ext_limb_angle(int k, int m) {
#pragma clang fp contract(off)
    if (k == m) return 3.14159265358979323846 - 1e-8;
    const double step = (3.14159265358979323846 - -3.14159265358979323846) / (double)m;
    const double y = (double)k * step;
    return y + -3.14159265358979323846;
}
// Source-plane limb point w0 + rho e^{i th} (extended_source.py:79-98), each component a
// product then a sum as the reference's complex arithmetic rounds it.
__device__ __forceinline__ cd ext_limb_point(cd w0, double rho, double cs, double sn) {
#pragma clang fp contract(off)
    const double x = rho * cs, y = rho * sn;
    return cd{x + w0.re, y + w0.im};
}

// ---- point solve -----------------------------------------------------------
// warm = true: z holds initial roots (those of a nearby limb point).  The solve ends
// with the compensated polish, which takes converged roots to the doubles nearest the
// exact roots, so the result does not depend on where the iteration started and the
// rows are put in order afterwards by value (match_points).  first = true (limb point
// 0): the row order of this point seeds the continuation of every row along the limb
// (greedy matching resolves competing images in row order), so it is solved by the
// reference-compatible FP64 form from the reference's initial estimates
// (init_est.h:57-102, ehrlich_aberth.h:60-150), which puts the roots in the reference
// solver's order, in both precision forms.
template <int NL, int MODE>
__device__ __forceinline__ bool ext_solve_point(cd w, const LensFrame<NL>& lf,
                                                cd (&z)[ext_degree<NL>()], uint32_t& flags,
                                                int itmax, SolveStats* stats = nullptr,
                                                bool warm = false, bool first = false) {
    constexpr int DEG = ext_degree<NL>();
    bool ok = true;
    flags = 0;
    if constexpr (NL == 1) {
        const double sq = sqrt(1.0 + 4.0 * rcp(norm2(w)));
        z[0] = (0.5 * (1.0 + sq)) * w;
        z[1] = (0.5 * (1.0 - sq)) * w;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            flags |= 1u << j;
            // sign of det J; the reference's 1 - 1/|z|^2 has the same sign as 1 - 1/|z|^4
            if (norm2(z[j]) < 1.0) flags |= 1u << (16 + j);
        }
    } else {
        if (first)
            ok = solve_lens_refcompat<NL>(w, lf, z, itmax, stats);
        else
            ok = solve_lens<NL, MODE, true>(w, lf, z, itmax, warm, stats);
#pragma unroll
        for (int j = 0; j < DEG; ++j) {
            ImageTerms<NL> it = image_terms<NL>(z[j], lf);
            if (norm2(z[j] - it.G - w) < 1e-12) flags |= 1u << j;
            if (1.0 - norm2(it.K) < 0.0) flags |= 1u << (16 + j);
        }
    }
    return ok;
}

template <int DEG>
__device__ __forceinline__ void load_pt(const cd* zb, int k, cd (&z)[DEG]) {
#pragma unroll
    for (int r = 0; r < DEG; ++r) z[r] = zb[(size_t)k * DEG + r];
}
template <int DEG>
__device__ __forceinline__ void store_pt(cd* zb, int k, const cd (&z)[DEG]) {
#pragma unroll
    for (int r = 0; r < DEG; ++r) zb[(size_t)k * DEG + r] = z[r];
}
template <int DEG>
__device__ __forceinline__ uint32_t permute_flags(uint32_t f, uint64_t p) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < DEG; ++i) {
        const int s = pget(p, i);
        o |= ((f >> s) & 1u) << i;
        o |= ((f >> (16 + s)) & 1u) << (16 + i);
    }
    return o;
}
template <int DEG>
__device__ __forceinline__ void permute_pt(cd (&z)[DEG], uint64_t p) {
    cd t[DEG];
#pragma unroll
    for (int i = 0; i < DEG; ++i) {
        cd v = z[0];
#pragma unroll
        for (int j = 1; j < DEG; ++j) v = (pget(p, i) == j) ? z[j] : v;
        t[i] = v;
    }
#pragma unroll
    for (int i = 0; i < DEG; ++i) z[i] = t[i];
}

// _permute_images over points 0..M-1 of buffer b (extended_source.py:34-53)
template <int DEG, int NT>
__device__ void chain_images(cd* zb, uint32_t* fb, uint64_t* perm, int M) {
    const int tid = threadIdx.x;
    for (int k = tid; k < M - 1; k += NT) {
        cd a[DEG], b[DEG];
        load_pt<DEG>(zb, k, a);
        load_pt<DEG>(zb, k + 1, b);
        perm[k] = nn_map<DEG>(a, b);
    }
    __syncthreads();
    if (tid < 64) {
        // wave 0, 64 points per batch: runs of steps without a conflict compose by a
        // wave-wide scan (composition is associative); a conflict step applies the
        // greedy rule to the carried rows, then the scan resumes after it.
        const int lane = tid;
        const uint64_t id = pident<DEG>();
        uint64_t pi = id;
        for (int base = 0; base < M; base += 64) {
            const int k_l = base + lane;
            const int nb = min(64, M - base);
            const bool valid = k_l < M - 1;
            const uint64_t e = valid ? perm[k_l] : id;
            const uint64_t cm = __ballot(valid && (e & kConflictBit));
            uint64_t mine = 0;
            int start = 0;
            while (start < nb) {
                const uint64_t rest = cm >> start;
                const int c = rest ? start + __ffsll((long long)rest) - 1 : nb;
                // x_j = step j for start <= j < c, identity elsewhere; inclusive scan
                uint64_t x = (lane >= start && lane < c) ? (e & ~kConflictBit) : id;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint64_t o = __shfl_up(x, off, 64);
                    if (lane >= off) x = pcompose<DEG>(x, o);
                }
                uint64_t ex = __shfl_up(x, 1, 64);
                if (lane == 0) ex = id;
                const int last = (c < nb) ? c : nb - 1;
                if (lane >= start && lane <= last) mine = pcompose<DEG>(ex, pi);
                const uint64_t Pc = __shfl(x, (c > start) ? c - 1 : 0, 64);
                const uint64_t pi_after = (c > start) ? pcompose<DEG>(Pc, pi) : pi;
                if (c < nb) {
                    // greedy match_points of the carried rows of point k against point
                    // k+1, one row per lane: lane i holds a_i, lane j holds b_j
                    const int k = base + c;
                    const cd av = (lane < DEG) ? zb[(size_t)k * DEG + pget(pi_after, lane)]
                                               : cd{0.0, 0.0};
                    const cd bv = (lane < DEG) ? zb[(size_t)(k + 1) * DEG + lane] : cd{0.0, 0.0};
                    unsigned used = 0;
                    uint64_t p = 0;
                    for (int i = 0; i < DEG; ++i) {
                        const cd ai = cd{__shfl(av.re, i, 64), __shfl(av.im, i, 64)};
                        double d = (lane < DEG && !((used >> lane) & 1u)) ? norm2(bv - ai)
                                                                          : INFINITY;
                        int bj = lane;
#pragma unroll
                        for (int off = 32; off >= 1; off >>= 1) {
                            const double od = __shfl_xor(d, off, 64);
                            const int oj = __shfl_xor(bj, off, 64);
                            if (od < d || (od == d && oj < bj)) {
                                d = od;
                                bj = oj;
                            }
                        }
                        used |= 1u << bj;
                        p |= (uint64_t)bj << (4 * i);
                    }
                    pi = p;
                    start = c + 1;
                } else {
                    pi = pi_after;
                    start = nb;
                }
            }
            if (k_l < M) perm[k_l] = mine;
        }
    }
    __syncthreads();
    for (int k = tid; k < M; k += NT) {
        // gather straight from memory (row i <- raw row p[i]): no register array with a
        // dynamic index, which the compiler would place in scratch
        const uint64_t p = perm[k];
        cd z[DEG];
#pragma unroll
        for (int i = 0; i < DEG; ++i) z[i] = zb[(size_t)k * DEG + pget(p, i)];
        store_pt<DEG>(zb, k, z);
        fb[k] = permute_flags<DEG>(fb[k], p);
    }
    __syncthreads();
}

// block-wide exclusive prefix sum of per-thread counts (NT threads)
template <int NT>
__device__ inline int block_exclusive_scan(int v, int* sh, int& total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        const int t = (tid >= off) ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    total = sh[NT - 1];
    const int r = sh[tid] - v;
    __syncthreads();
    return r;
}

__device__ inline double block_sum(double v, double* sh) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kExtBlock / 64; ++q) s += sh[q];
    __syncthreads();
    return s;
}

// ---- segments and contours ---------------------------------------------------
struct Seg {
    int row, start, len;  // points start..start+len-1 of image row `row`
    int par;              // parity of the segment (+1/-1)
    int empty;
    double length;        // sum of |z_{k+1} - z_k|
    cd head, tail;        // first and last point (0 for an empty segment)
};
struct Piece {
    int row, start, len, rev;
};
struct Contour {
    Piece pc[kExtMaxPieces];
    int np;
    int npts;  // points (without the closing point)
    int par;   // parity of the first point
};

template <int DEG>
__device__ __forceinline__ cd pt(const cd* zb, int k, int row) {
    return zb[(size_t)k * DEG + row];
}
// point k of a piece list
template <int DEG>
__device__ inline cd contour_point(const cd* zb, const Piece* pc, int np, int k) {
    for (int i = 0; i < np; ++i) {
        if (k < pc[i].len) {
            const int idx = pc[i].rev ? pc[i].start + pc[i].len - 1 - k : pc[i].start + k;
            return pt<DEG>(zb, idx, pc[i].row);
        }
        k -= pc[i].len;
    }
    return cd{0.0, 0.0};
}

}  // namespace caus

namespace caus {

// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n (ascending x)
__device__ inline void gauss_legendre(int n, double* x, double* w) {
    for (int i = threadIdx.x; i < n; i += kExtBlock) {
        // root i counted from x = +1 downwards, stored ascending
        double z = cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
        double pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p1 = 1.0, p2 = 0.0;
        for (int j = 1; j <= n; ++j) {
            const double p3 = p2;
            p2 = p1;
            p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

// sqrt(x) for x > 0 in the normal range (x >= 2^-767), with inv ~ 1/sqrt(x): the
// v_rsq_f64 seed, one Newton step on h ~ 1/(2 sqrt x) and two residual corrections of
// g ~ sqrt x (the same sequence as the compiler's sqrt, which adds range scaling and
// special cases that these callers do not need), then one more step on h for inv.
__device__ __forceinline__ double sqrt_rsq(double x, double& inv) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    inv = 2.0 * fma(h, fma(-h, g, 0.5), h);
    return g;
}

// Brightness profile at z (integrate.py:29-44): r = |lens_eq(z) - w0| / rho,
// B = c3 (u1 B_r + 1 - 2 u1) with c3 = 3/(3-u1), B_r = 1 + sqrt(1 - r^2) (r <= 1) or
// 1 - sqrt(1 - 1/r^2).  Also returned for the adjoint: v_j = 1/(conj z - conj r_j),
// D = lens_eq(z) - w0, 1/|D| and dB_r/dr.  Contraction is off here so that the
// forward-only and the forward+VJP kernels, which inline this into different code, round
// B alike.
template <int NL>
__device__ __forceinline__ double ld_brightness_core(cd z, const LensFrame<NL>& lf, cd w0,
                                                     double rho, double irho, double u1,
                                                     double c3, cd (&v)[NL], cd& D,
                                                     double& iaD, double& r, double& Br,
                                                     double& dBr) {
#pragma clang fp contract(off)
    const cd zb = conj(z);
    // v_j = 1/(conj z - conj r_j): the squared moduli share one reciprocal of their product
    cd d[NL];
    double n[NL], sj[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        d[j] = zb - conj(lf.r[j]);
        n[j] = norm2(d[j]);
    }
    if constexpr (NL == 1) {
        sj[0] = rcp(n[0]);
    } else if constexpr (NL == 2) {
        const double rP = rcp(n[0] * n[1]);
        sj[0] = n[1] * rP;
        sj[1] = n[0] * rP;
    } else {
        const double n01 = n[0] * n[1];
        const double rP = rcp(n01 * n[2]);
        const double t = n[2] * rP;
        sj[0] = n[1] * t;
        sj[1] = n[0] * t;
        sj[2] = n01 * rP;
    }
    cd G = cd{0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        v[j] = cd{d[j].re * sj[j], -d[j].im * sj[j]};
        G = cd{fma(lf.e[j], v[j].re, G.re), fma(lf.e[j], v[j].im, G.im)};  // explicit: no
    }                                                                       // contraction choice
    D = z - G - w0;
    const double n2 = norm2(D);
    double ia;
    const double aD = sqrt_rsq(n2, ia);
    const bool pos = n2 > 0.0;
    iaD = pos ? ia : 0.0;
    r = pos ? aD * irho : 0.0;
    const bool inside = r <= 1.0;
    const double q = inside ? r : rho * iaD;   // 1/r = rho / |D|
    const double x = 1.0 - q * q;
    double is;
    const double sq = sqrt_rsq(x, is);
    const bool xp = x > 0.0;
    const double s = xp ? sq : 0.0;
    is = xp ? is : 0.0;
    Br = inside ? 1.0 + s : 1.0 - s;
    dBr = inside ? -r * is : -(q * q * q) * is;
    return c3 * (u1 * Br + 1.0 - 2.0 * u1);
}

// c3 = 3/(3-u1), as the brightness uses it
__device__ __forceinline__ double ld_c3(double u1) { return 3.0 / (3.0 - u1); }

template <int NL>
__device__ __forceinline__ double ld_brightness(cd z, const LensFrame<NL>& lf, cd w0, double rho,
                                                double irho, double u1, double c3) {
    cd v[NL], D;
    double iaD, r, Br, dBr;
    return ld_brightness_core<NL>(z, lf, w0, rho, irho, u1, c3, v, D, iaD, r, Br, dBr);
}

// one trapezoid term of the limb-darkened contour integral (integrate.py:118-121),
// written with explicit fma so that both kernels that form it round it alike
__device__ __forceinline__ double ld_trapezoid(double Pk, double Pn, double Qk, double Qn, cd a,
                                               cd b, double acc) {
    return fma(0.5 * (Pk + Pn), b.re - a.re, fma(0.5 * (Qk + Qn), b.im - a.im, acc));
}

struct GLNodes {
    const double *x1, *w1, *x2, *w2;
    int n1, n2;
};

// -0.5 * int_{a}^{b} B(x + i y) dy  (P) or 0.5 * int_{a}^{b} B(x + i y) dx (Q), split at
// the reference's point (integrate.py:50-107)
template <int NL>
__device__ __forceinline__ double ld_line(bool isP, double fixed, double a, double b, double rho,
                                          const LensFrame<NL>& lf, cd w0, double u1,
                                          const GLNodes& g) {
    const double irho = rcp(rho);
    const double d = fabs(b - a);
    // explicit fma: the forward+VJP twin (ld_line_adj) must round the split point alike
    double sp = (b > a) ? fma(-2.0, rho, b) : fma(2.0, rho, b);
    if (0.5 * d <= 2.0 * rho) sp = fma(0.5, d, a);
    const double c3 = ld_c3(u1);
    double I = 0.0;
    for (int part = 0; part < 2; ++part) {
        const double lo = part == 0 ? a : sp, hi = part == 0 ? sp : b;
        const double* x = part == 0 ? g.x1 : g.x2;
        const double* wt = part == 0 ? g.w1 : g.w2;
        const int n = part == 0 ? g.n1 : g.n2;
        const double h = 0.5 * (hi - lo), c = 0.5 * (hi + lo);
        double s = 0.0;
#pragma unroll 2
        for (int i = 0; i < n; ++i) {
            const double v = fma(h, x[i], c);
            const cd zz = isP ? cd{fixed, v} : cd{v, fixed};
            s = fma(h * ld_brightness<NL>(zz, lf, w0, rho, irho, u1, c3), wt[i], s);
        }
        I += s;
    }
    return isP ? -0.5 * I : 0.5 * I;
}

// Uniform-source contour integral over points c_0..c_{L-1} (pieces), closed back to
// c_0: nterms trapezoid terms of (x dy - y dx)/2 (L, or L-1 for the single-lens quirk).
template <int NL, int DEG>
__device__ double integrate_contour(const cd* zb, const Piece* pc, int np, int L, int nterms,
                                    double* shd) {
    const int tid = threadIdx.x;
    double acc = 0.0;
    if (L <= 0) return block_sum(0.0, shd);
    for (int k = tid; k < nterms; k += kExtBlock) {
        const cd a = contour_point<DEG>(zb, pc, np, k);
        const cd b = contour_point<DEG>(zb, pc, np, (k + 1 == L) ? 0 : k + 1);
        acc += 0.5 * (0.5 * a.re + 0.5 * b.re) * (b.im - a.im) +
               0.5 * (-0.5 * a.im - 0.5 * b.im) * (b.re - a.re);
    }
    return block_sum(acc, shd);
}

// ---- stitching (extended_source.py:312-667), wave-uniform code on wave 0 ---------
struct ActSeg {
    Piece pc[kExtMaxPieces];
    int np, t, par, empty;
};

template <int DEG>
__device__ inline cd seg_point(const cd* zb, const Seg& s, int k) {
    if (s.empty || k < 0 || k >= s.len) return cd{0.0, 0.0};
    return pt<DEG>(zb, s.start + k, s.row);
}
template <int DEG>
__device__ inline cd act_point(const cd* zb, const ActSeg& a, int k) {
    if (a.empty || k < 0 || k > a.t) return cd{0.0, 0.0};
    return contour_point<DEG>(zb, a.pc, a.np, k);
}

// line ending at the head / tail (second element is the connection point).
// Dense-array semantics: points beyond the segment are zero, dynamic_slice clamps.
template <class F>
__device__ inline void head_line(F get, int t, cd& l0, cd& l1) {
    const cd x0 = get(0), x1 = get(1), x2 = get(2);
    const bool c = (norm2(x1 - x0) > 1e-10) || (t <= 1);
    if (c) {
        l0 = x1;
        l1 = x0;
    } else {
        l0 = x2;
        l1 = x1;
    }
}
template <class F>
__device__ inline void tail_line(F get, int t, int dense_len, cd& l0, cd& l1) {
    const bool c = (t <= 1) || (norm2(get(t) - get(t - 1)) > 1e-10);
    int s = c ? t - 1 : t - 2;
    s = s < 0 ? 0 : (s > dense_len - 2 ? dense_len - 2 : s);
    l0 = get(s);
    l1 = get(s + 1);
}

__device__ inline bool geom_condition(cd a0, cd a1, cd b0, cd b1) {
    const double dist = sqrt(norm2(a1 - b1));
    const bool c1 = dist < 1e-1;
    const cd v1 = a1 - a0, v2 = b1 - b0;
    const double n1 = sqrt(norm2(v1)), n2 = sqrt(norm2(v2));
    const double cosang = (v1.re / n1) * (v2.re / n2) + (v1.im / n1) * (v2.im / n2);
    const double alpha = acos(cosang);  // NaN outside [-1,1] or for a zero vector
    const bool c2 = (180.0 - alpha * (180.0 / 3.14159265358979323846)) < 60.0;
    const bool c3 = sqrt(norm2(a1 - b1)) < sqrt(norm2(a0 - b0));
    return (c1 && c2 && c3) || dist < 1e-5;
}

}  // namespace caus

// ============================================================================
// Reverse mode of the contour pipeline.  Directions (user order):
//   0 Re w0, 1 Im w0, 2 rho, 3 u1, 4.. lens parameters (s, q[, q3, r3, psi]).
// Discrete choices (limb angles, rows, segment structure) are piecewise constant;
// every contour point is the image z of a fixed limb angle, differentiated
// implicitly through the lens equation, and the Gauss-Legendre line integrals are
// differentiated exactly as discretised (moving end points and split point included).
// ============================================================================
namespace caus {

template <int NL>
constexpr int ext_nd() {
    return 4 + (NL == 1 ? 0 : n_user_params<NL>());
}

template <int NL>
struct ExtTangents {
    static constexpr int ND = ext_nd<NL>();
    cd dw0[ND];              // tangent of the lens-frame source centre
    LensTangent<NL> lt[ND];  // tangent of the lens frame
    double drho[ND], du1[ND];
};

template <int NL>
__device__ inline void ext_tangents(const double* prm, ExtTangents<NL>& T) {
    constexpr int ND = ext_nd<NL>();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            T.lt[d].de[j] = 0.0;
            T.lt[d].dr[j] = cd{0.0, 0.0};
        }
        T.lt[d].dxcm = 0.0;
        T.dw0[d] = cd{0.0, 0.0};
        T.drho[d] = 0.0;
        T.du1[d] = 0.0;
    }
    T.dw0[0] = cd{1.0, 0.0};
    T.dw0[1] = cd{0.0, 1.0};
    T.drho[2] = 1.0;
    T.du1[3] = 1.0;
    if constexpr (NL > 1) {
        LensTangent<NL> up[n_user_params<NL>()];
        lens_frame_tangents<NL>(prm, up);
#pragma unroll
        for (int k = 0; k < n_user_params<NL>(); ++k) {
            T.lt[4 + k] = up[k];
            T.dw0[4 + k] = cd{up[k].dxcm, 0.0};
        }
    }
}

// point k of a piece list: value, limb index and row
template <int DEG>
__device__ inline cd contour_locate(const cd* zb, const Piece* pc, int np, int k, int& idx,
                                    int& row) {
    for (int i = 0; i < np; ++i) {
        if (k < pc[i].len) {
            idx = pc[i].rev ? pc[i].start + pc[i].len - 1 - k : pc[i].start + k;
            row = pc[i].row;
            return pt<DEG>(zb, idx, row);
        }
        k -= pc[i].len;
    }
    idx = 0;
    row = 0;
    return cd{0.0, 0.0};
}

// Reverse-mode accumulators of the direct (not through the contour points) dependence
// of the limb-darkened line integrals on the lens frame, the source centre, rho and u1.
// A direction with tangents (de_j, dr_j, dw0, drho, du1) receives
//   sum_j de_j ge_j + e_j Re(gr_j conj(dr_j)) + Re(gw conj(dw0)) + grho drho + gu1 du1
// (the mass e_j of the dr_j term is applied once, in along), so that each
// Gauss-Legendre node costs O(NL) instead of O(NL x directions).
template <int NL>
struct LDAdj {
    double ge[NL];
    cd gr[NL];
    cd gw;
    double grho, gu1;
    __device__ void zero() {
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            ge[j] = 0.0;
            gr[j] = cd{0.0, 0.0};
        }
        gw = cd{0.0, 0.0};
        grho = gu1 = 0.0;
    }
    __device__ void axpy(double s, const LDAdj& o) {
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            ge[j] += s * o.ge[j];
            gr[j] = gr[j] + s * o.gr[j];
        }
        gw = gw + s * o.gw;
        grho += s * o.grho;
        gu1 += s * o.gu1;
    }
    __device__ double along(const LensTangent<NL>& lt, const double (&e)[NL], cd dw0, double drho,
                            double du1) const {
        double v = grho * drho + gu1 * du1 + gw.re * dw0.re + gw.im * dw0.im;
#pragma unroll
        for (int j = 0; j < NL; ++j)
            v += lt.de[j] * ge[j] + e[j] * (gr[j].re * lt.dr[j].re + gr[j].im * lt.dr[j].im);
        return v;
    }
};

// brightness B at z (integrate.py:29-44) and its derivatives along x and y at fixed
// parameters (Bx, By); c * (direct parameter derivative of B) is added to A.
// r = |D|/rho with D = lens_eq(z) - w0;  dr = Re(conj(D) dD)/(|D| rho) - r drho/rho,
// dD = -(dG + dw0),  dG = sum_j de_j v_j + e_j v_j^2 conj(dr_j),  v_j = 1/(conj z - conj r_j).
// dc3 = dc3/du1 = c3^2/3.
template <int NL>
__device__ inline double ld_brightness_adj(cd z, const LensFrame<NL>& lf, cd w0, double rho,
                                           double irho, double u1, double c3, double dc3, double c,
                                           double& Bx, double& By, LDAdj<NL>& A) {
    cd v[NL], D;
    double iaD, r, Br, dBr;
    const double B = ld_brightness_core<NL>(z, lf, w0, rho, irho, u1, c3, v, D, iaD, r, Br, dBr);
    cd vv[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) vv[j] = v[j] * v[j];
    cd K = lf.e[0] * vv[0];
#pragma unroll
    for (int j = 1; j < NL; ++j) K = K + lf.e[j] * vv[j];
    const double dBdr = c3 * u1 * dBr;
    const double iDr = iaD * irho;
    // dD/dx = 1 + K, dD/dy = i (1 - K)
    const double gx = D.re * (1.0 + K.re) + D.im * K.im;
    const double gy = D.re * K.im + D.im * (1.0 - K.re);
    const double f = dBdr * iDr;
    Bx = f * gx;
    By = f * gy;
    const double dBdu1 = dc3 * (u1 * Br + 1.0 - 2.0 * u1) + c3 * (Br - 2.0);
    const double t = c * dBdr * irho;
    const double a = t * iaD;
    const cd Dc = conj(D);
    A.gw = A.gw - a * D;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        A.ge[j] -= a * (D.re * v[j].re + D.im * v[j].im);
        A.gr[j] = A.gr[j] - a * (Dc * vv[j]);
    }
    A.grho -= t * r;
    A.gu1 += c * dBdu1;
    return B;
}

// P (isP) or Q line integral with its partial derivatives along the fixed coordinate,
// the lower limit a and the upper limit b; wA * (direct parameter derivative) goes to A.
template <int NL>
__device__ inline double ld_line_adj(bool isP, double fixed, double a, double b, double rho,
                                     const LensFrame<NL>& lf, cd w0, double u1, const GLNodes& g,
                                     double wA, double& dfix, double& da, double& db,
                                     LDAdj<NL>& A) {
    const double d = fabs(b - a);
    const double irho = rcp(rho);
    const double sc = isP ? -0.5 : 0.5;
    double sp, dsp_da, dsp_db, dsp_drho;
    if (0.5 * d <= 2.0 * rho) {
        const double sg = (b - a) > 0 ? 1.0 : ((b - a) < 0 ? -1.0 : 0.0);
        sp = fma(0.5, d, a);
        dsp_da = 1.0 - 0.5 * sg;
        dsp_db = 0.5 * sg;
        dsp_drho = 0.0;
    } else if (b > a) {
        sp = fma(-2.0, rho, b);
        dsp_da = 0.0;
        dsp_db = 1.0;
        dsp_drho = -2.0;
    } else {
        sp = fma(2.0, rho, b);
        dsp_da = 0.0;
        dsp_db = 1.0;
        dsp_drho = 2.0;
    }
    const double c3 = ld_c3(u1), dc3 = c3 * c3 / 3.0;
    double S = 0.0, dS_lo[2], dS_hi[2];
    dfix = 0.0;
    for (int part = 0; part < 2; ++part) {
        const double lo = part == 0 ? a : sp, hi = part == 0 ? sp : b;
        const double* x = part == 0 ? g.x1 : g.x2;
        const double* wt = part == 0 ? g.w1 : g.w2;
        const int n = part == 0 ? g.n1 : g.n2;
        const double h = 0.5 * (hi - lo), c = 0.5 * (hi + lo);
        const double cw = wA * sc * h;
        // s: the forward sum, rounded as ld_line rounds it; S0 = sum w B,
        // S1 = sum w dB/dt, S2 = sum w x dB/dt, S3 = sum w dB/d(fixed) give the
        // derivatives along lo, hi and the fixed coordinate
        double s = 0.0, S0 = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0;
#pragma unroll 1
        for (int i = 0; i < n; ++i) {
            const double xi = x[i], wi = wt[i];
            const double t = fma(h, xi, c);
            const cd zz = isP ? cd{fixed, t} : cd{t, fixed};
            double Bx, By;
            const double B = ld_brightness_adj<NL>(zz, lf, w0, rho, irho, u1, c3, dc3, cw * wi, Bx,
                                                   By, A);
            const double fl = isP ? By : Bx;   // along the line
            const double ff = isP ? Bx : By;   // across (fixed coordinate)
            s = fma(h * B, wi, s);
            S0 = fma(wi, B, S0);
            const double wfl = wi * fl;
            S1 += wfl;
            S2 = fma(wfl, xi, S2);
            S3 = fma(wi, ff, S3);
        }
        S += s;
        // d/dlo of h sum w B(c + h x) with dh/dlo = -1/2, dc/dlo = 1/2, and d/dhi
        dS_lo[part] = fma(0.5 * h, S1 - S2, -0.5 * S0);
        dS_hi[part] = fma(0.5 * h, S1 + S2, 0.5 * S0);
        dfix = fma(h, S3, dfix);
    }
    const double dsp = dS_hi[0] + dS_lo[1];
    da = dS_lo[0] + dsp * dsp_da;
    db = dS_hi[1] + dsp * dsp_db;
    A.grho += wA * sc * dsp * dsp_drho;
    dfix *= sc;
    da *= sc;
    db *= sc;
    return sc * S;
}

// Adjoint of one contour integral: accumulates scale * dI/dtheta into g[ND] (per thread).
// pq holds P, Q of the contour points (limb darkening; written by the forward pass).
// Adjoint of the image z of limb angle th for the cotangent zbar = (bx, by) of z:
// dz = (u - K conj u)/det with u = dw0 + drho e^{i th} + sum_j de_j v_j + e_j v_j^2 conj(dr_j)
// gives zbar . dz = Re(c u), c = (conj(zbar) - zbar conj(K))/det, added to A.
template <int NL>
__device__ inline void image_adjoint(cd z, double th, const LensFrame<NL>& lf, double bx,
                                     double by, LDAdj<NL>& A) {
    const ImageTerms<NL> it = image_terms<NL>(z, lf);
    const double idet = rcp(1.0 - norm2(it.K));
    const cd zbar = cd{bx, by};
    const cd c = idet * (conj(zbar) - zbar * conj(it.K));
    double sn, cs;
    sincos(th, &sn, &cs);
    A.gw = A.gw + conj(c);
    A.grho += c.re * cs - c.im * sn;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const cd v = it.inv[j];
        const cd cv = c * v;
        A.ge[j] += cv.re;
        A.gr[j] = A.gr[j] + cv * v;
    }
}

// Adjoint of one uniform-source contour integral: adds scale * dI/d(frame, w0, rho) to
// adj (per thread).  Each point's cotangent, from the two trapezoid terms it enters, is
// pushed through the image tangents.
template <int NL, int DEG>
__device__ void contour_adjoint(const cd* zb, const double* theta, const Piece* pc, int np, int L,
                                int nterms, const LensFrame<NL>& lf, double scale,
                                LDAdj<NL>& adj) {
    const int tid = threadIdx.x;
    for (int k = tid; k < L; k += kExtBlock) {
        int idx, row;
        const cd zk = contour_locate<DEG>(zb, pc, np, k, idx, row);
        const int kn = (k + 1 == L) ? 0 : k + 1;
        const int kp = (k == 0) ? L - 1 : k - 1;
        const cd zn = contour_point<DEG>(zb, pc, np, kn);
        const cd zp = contour_point<DEG>(zb, pc, np, kp);
        double xbar = 0.0, ybar = 0.0;
        if (k < nterms) {
            xbar += 0.5 * zn.im;
            ybar -= 0.5 * zn.re;
        }
        if (kp < nterms) {
            xbar -= 0.5 * zp.im;
            ybar += 0.5 * zp.re;
        }
        image_adjoint<NL>(zk, theta[idx], lf, scale * xbar, scale * ybar, adj);
    }
}

// One contour of a source as the integrals see it: its pieces, L points, nterms
// trapezoid terms, the divisor of its centroid sum, its parity, the offset of its
// points in the source's flat point list and (limb darkening) its centroid.
struct ExtCEntry {
    const Piece* pc;
    int np, L, nterms, off;
    double cdiv, par, zx, zy;
};
constexpr int kExtMaxContours = 16;  // >= nlenses^2 + 1 closed rows + 3 stitched contours

// Limb-darkened integrals of all nc contours of a source (integrate.py:47-121), their
// parity-weighted sum, and with VJP the reverse mode (adds to adj).  The P and Q line
// integrals of every point -- 2 npts_ld brightness evaluations per point, nearly all of
// the work -- run as one flat loop over the source's T points, so the workgroup idles only
// in its last iteration; the centroids before it and the trapezoid sums and the point
// adjoints after it run contour by contour, and every sum rounds as it would in a
// contour-by-contour evaluation.  The forward and forward+VJP kernels form P and Q
// alike (ld_line / ld_line_adj), so they return the same integral.
// pq, xy, yz hold 2 T doubles each: P and Q, the points' cotangents, and the parts of the
// centroid's cotangent (Pbar dP/dy0, Qbar dQ/dx0).
template <int NL, int DEG, bool VJP>
__device__ double integrate_source_ld(const cd* zb, const double* theta, ExtCEntry* ct, int nc,
                                      int T, double* pq, double* xy, double* yz,
                                      const LensFrame<NL>& lf, cd w0, double rho, double u1,
                                      const GLNodes& gl, LDAdj<NL>& adj, double* shd) {
    const int tid = threadIdx.x;
    // centroids (integrate.py:112); the closing point repeats c_0 in the sum
    for (int c = 0; c < nc; ++c) {
        const ExtCEntry& e = ct[c];
        double sx = 0.0, sy = 0.0;
        for (int k = tid; k < e.L; k += kExtBlock) {
            const cd a = contour_point<DEG>(zb, e.pc, e.np, k);
            sx += a.re;
            sy += a.im;
        }
        const cd c0 = (e.L > 0) ? contour_point<DEG>(zb, e.pc, e.np, 0) : cd{0.0, 0.0};
        const double zx = (block_sum(sx, shd) + c0.re) / e.cdiv;
        const double zy = (block_sum(sy, shd) + c0.im) / e.cdiv;
        if (tid == 0) {
            ct[c].zx = zx;
            ct[c].zy = zy;
        }
    }
    __syncthreads();
    // the P and Q lines of every point, flat over the contours
    for (int gi = tid; gi < T; gi += kExtBlock) {
        int c = 0;
        while (c + 1 < nc && gi >= ct[c + 1].off) ++c;
        const ExtCEntry& e = ct[c];
        const int k = gi - e.off;
        const cd zk = contour_point<DEG>(zb, e.pc, e.np, k);
        if constexpr (!VJP) {
            pq[2 * gi] = ld_line<NL>(true, zk.re, e.zy, zk.im, rho, lf, w0, u1, gl);
            pq[2 * gi + 1] = ld_line<NL>(false, zk.im, e.zx, zk.re, rho, lf, w0, u1, gl);
        } else {
            const int L = e.L;
            const int kn = (k + 1 == L) ? 0 : k + 1;
            const int kp = (k == 0) ? L - 1 : k - 1;
            const cd zn = contour_point<DEG>(zb, e.pc, e.np, kn);
            const cd zp = contour_point<DEG>(zb, e.pc, e.np, kp);
            // cotangents of P_k, Q_k from the two trapezoid terms they enter
            double Pbar = 0.0, Qbar = 0.0;
            if (k < e.nterms) {
                Pbar += 0.5 * (zn.re - zk.re);
                Qbar += 0.5 * (zn.im - zk.im);
            }
            if (kp < e.nterms) {
                Pbar += 0.5 * (zk.re - zp.re);
                Qbar += 0.5 * (zk.im - zp.im);
            }
            double dfP, daP, dbP, dfQ, daQ, dbQ;
            pq[2 * gi] = ld_line_adj<NL>(true, zk.re, e.zy, zk.im, rho, lf, w0, u1, gl,
                                         e.par * Pbar, dfP, daP, dbP, adj);
            pq[2 * gi + 1] = ld_line_adj<NL>(false, zk.im, e.zx, zk.re, rho, lf, w0, u1, gl,
                                             e.par * Qbar, dfQ, daQ, dbQ, adj);
            xy[2 * gi] = Pbar * dfP + Qbar * dbQ;
            xy[2 * gi + 1] = Pbar * dbP + Qbar * dfQ;
            yz[2 * gi] = Pbar * daP;       // P_k = P(x_k; y0 .. y_k): to the centroid's y
            yz[2 * gi + 1] = Qbar * daQ;   // Q_k = Q(y_k; x0 .. x_k): to the centroid's x
        }
    }
    __syncthreads();
    double total = 0.0;
    for (int c = 0; c < nc; ++c) {
        const ExtCEntry& e = ct[c];
        const int L = e.L;
        const double* q = pq + 2 * e.off;
        double* v = xy + 2 * e.off;
        const double* y = yz + 2 * e.off;
        double acc = 0.0, x0bar = 0.0, y0bar = 0.0;
        for (int k = tid; k < L; k += kExtBlock) {
            const int kn = (k + 1 == L) ? 0 : k + 1;
            if (k < e.nterms) {
                const cd a = contour_point<DEG>(zb, e.pc, e.np, k);
                const cd b = contour_point<DEG>(zb, e.pc, e.np, kn);
                acc = ld_trapezoid(q[2 * k], q[2 * kn], q[2 * k + 1], q[2 * kn + 1], a, b, acc);
            }
            if constexpr (VJP) {
                const int kp = (k == 0) ? L - 1 : k - 1;
                double xbar = v[2 * k], ybar = v[2 * k + 1];
                if (k < e.nterms) {
                    xbar -= 0.5 * (q[2 * k] + q[2 * kn]);
                    ybar -= 0.5 * (q[2 * k + 1] + q[2 * kn + 1]);
                }
                if (kp < e.nterms) {
                    xbar += 0.5 * (q[2 * kp] + q[2 * k]);
                    ybar += 0.5 * (q[2 * kp + 1] + q[2 * k + 1]);
                }
                v[2 * k] = xbar;
                v[2 * k + 1] = ybar;
                y0bar += y[2 * k];
                x0bar += y[2 * k + 1];
            }
        }
        total += e.par * block_sum(acc, shd);
        if constexpr (VJP) {
            const double X0 = block_sum(x0bar, shd) / e.cdiv;
            const double Y0 = block_sum(y0bar, shd) / e.cdiv;
            // each thread reads back only the cotangents it wrote above
            for (int k = tid; k < L; k += kExtBlock) {
                int idx, row;
                const cd zk = contour_locate<DEG>(zb, e.pc, e.np, k, idx, row);
                const double wgt = (k == 0) ? 2.0 : 1.0;  // c_0 enters the centroid twice
                image_adjoint<NL>(zk, theta[idx], lf, e.par * (v[2 * k] + wgt * X0),
                                  e.par * (v[2 * k + 1] + wgt * Y0), adj);
            }
        }
    }
    return total;
}

}  // namespace caus

// ============================================================================
// Batch pipeline.  The contours of a batch of sources are computed by a fixed
// sequence of launches over the whole batch, so that the root solves (nearly all of
// the arithmetic) run one limb point per lane at the solver's occupancy:
//   k_ext_plan        1 block      effective npts, prefix offsets of the flat work lists
//   k_ext_limb_init   flat         initial limb angles, kExtPer consecutive angles per lane
//   k_ext_chain       per source   image continuation (extended_source.py:34-53), limb
//                                  order and gap keys of the initial points
//   10 x k_ext_select per source   gap-key sort, insertion shifts, next limb order, tasks
//        k_ext_solve_add flat      new points appended, matched to their left neighbour,
//                                  keys of their two gaps
//   k_ext_chain       per source   points gathered in limb order, image continuation
//   k_ext_contours    per source   closure test, segments, stitching -> piece lists
//   k_ext_integrate   per source   contour integrals, magnification, reverse mode
// No launch allocates, synchronises with the host or clears memory with a memset, so
// the sequence can be captured in a HIP graph.
// ============================================================================
namespace caus {

// waves per SIMD that the register allocation of the reverse-mode integrate kernel
// targets (its per-thread adjoint state is large)
#ifndef CAUS_EXT_VJP_WAVES
#define CAUS_EXT_VJP_WAVES 3
#endif

constexpr int kExtPer = 4;          // initial limb angles per lane of k_ext_limb_init
constexpr int kExtFlatBlock = 256;  // block of the flat kernels
constexpr int kExtPlanBlock = 1024;

struct ExtArgs {
    const cd* w0;
    int64_t S;
    const double* rho;
    const double* params;
    int64_t pstride;
    const int32_t* npts;
    int npts_def;
    const double* u1;
    int64_t u1stride;
    int npts_ld;
    double* mag;
    uint8_t* status;
    const double* g_mag;   // VJP only
    double* g_out;         // VJP only: (S, 4 + P)
    char* ws;              // batch header, then one block of ws_per_source bytes per source
    size_t ws_per_source;
    unsigned long long* work;  // optional counters: FP64 root-iterations, FP32
                               // root-iterations, compensated root-iterations, limb-point
                               // solves, contour points
    int mmax;
    int nadd_max;
    int itmax;
};

// contours of one source, written by k_ext_contours
struct ExtContours {
    unsigned closed;  // image rows that are closed contours by themselves
    int open;         // 1 when the three stitched contours in con[] are used
    ActSeg con[3];
};

// batch header of the workspace
struct ExtBatch {
    int64_t* off_init;  // S + 1: prefix of the lanes of k_ext_limb_init
    int64_t* off_add;   // S + 1: prefix of the new points of one refinement round
    int32_t* npts;      // S: effective limb resolution
    uint8_t* fail;      // S: a solve did not converge, or npts was out of range
    int32_t* t_src;     // new-point tasks of the current round (capacity S * nadd_max):
    int32_t* t_left;    // source, append index of the left and right limb neighbours,
    int32_t* t_right;   // limb-order position and append index of the new point, and
    int32_t* t_dst;     // its limb angle
    int32_t* t_app;
    double* t_th;
};

__host__ __device__ inline size_t ext_task_cap(int64_t S, int nadd_max) {
    return (size_t)S * (size_t)(nadd_max > 0 ? nadd_max : 1);
}

__host__ __device__ inline size_t ext_header_bytes(int64_t S, int nadd_max) {
    const size_t cap = ext_task_cap(S, nadd_max);
    return ext_align(2 * (size_t)(S + 1) * sizeof(int64_t)) + ext_align((size_t)S * 4) +
           ext_align((size_t)S) + ext_align(5 * cap * sizeof(int32_t)) +
           ext_align(cap * sizeof(double));
}

__host__ __device__ inline size_t ext_source_bytes(int deg, int mmax) {
    return ext_ws_bytes(deg, mmax) + ext_align(sizeof(ExtContours));
}

__host__ __device__ inline size_t ext_total_bytes(int deg, int64_t S, int npts_max) {
    return ext_header_bytes(S, ext_n_add(npts_max)) +
           (size_t)S * ext_source_bytes(deg, ext_mmax(npts_max));
}

__device__ inline ExtBatch ext_batch(const ExtArgs& A) {
    ExtBatch b;
    char* p = A.ws;
    const size_t cap = ext_task_cap(A.S, A.nadd_max);
    b.off_init = (int64_t*)p;
    b.off_add = b.off_init + (A.S + 1);
    p += ext_align(2 * (size_t)(A.S + 1) * sizeof(int64_t));
    b.npts = (int32_t*)p;
    p += ext_align((size_t)A.S * 4);
    b.fail = (uint8_t*)p;
    p += ext_align((size_t)A.S);
    b.t_src = (int32_t*)p;
    b.t_left = b.t_src + cap;
    b.t_right = b.t_left + cap;
    b.t_dst = b.t_right + cap;
    b.t_app = b.t_dst + cap;
    p += ext_align(5 * cap * sizeof(int32_t));
    b.t_th = (double*)p;
    return b;
}

template <int DEG>
__device__ inline char* ext_source_base(const ExtArgs& A, int64_t s) {
    return A.ws + ext_header_bytes(A.S, A.nadd_max) + (size_t)s * A.ws_per_source;
}
template <int DEG>
__device__ inline ExtContours* ext_contours(const ExtArgs& A, int64_t s) {
    return (ExtContours*)(ext_source_base<DEG>(A, s) + ext_ws_bytes(DEG, A.mmax));
}

// lens frame, lens-frame source centre and radius of source s
template <int NL>
__device__ inline void ext_source(const ExtArgs& A, int64_t s, LensFrame<NL>& lf, cd& w0,
                                  double& rho) {
    double x_cm = 0.0;
    if constexpr (NL == 1) {
        lf.e[0] = 1.0;
        lf.r[0] = cd{0.0, 0.0};
    } else {
        lens_frame<NL>(A.params + s * A.pstride, lf, x_cm);
    }
    w0 = A.w0[s] + cd{x_cm, 0.0};
    rho = A.rho[s];
}

__device__ inline void ext_count(unsigned long long* work, const SolveStats& st, int solves) {
    atomicAdd(work + 0, (unsigned long long)st.it64);
    atomicAdd(work + 1, (unsigned long long)st.it32);
    atomicAdd(work + 2, (unsigned long long)st.itcomp);
    atomicAdd(work + 3, (unsigned long long)solves);
}

// limb points of source s after `round` refinement rounds
__device__ inline int ext_points(int npts, int round) {
    return ext_npts_init(npts) + round * ext_n_add(npts);
}

constexpr double kExtPi = 3.14159265358979323846;

// ---- plan: effective npts (out-of-range values clamped and flagged) and the prefix
// offsets of the flat work lists, one block scanning the batch in chunks
static __global__ __launch_bounds__(kExtPlanBlock) void k_ext_plan(ExtArgs A) {
    __shared__ int64_t s_a[kExtPlanBlock], s_b[kExtPlanBlock];
    const ExtBatch B = ext_batch(A);
    const int tid = threadIdx.x;
    int64_t carry_a = 0, carry_b = 0;
    for (int64_t base = 0; base < A.S; base += kExtPlanBlock) {
        const int64_t s = base + tid;
        int64_t ca = 0, cb = 0;
        if (s < A.S) {
            int n = A.npts ? A.npts[s] : A.npts_def;
            bool bad = false;
            if (n != 0 && (n < 4 || n > A.npts_def)) {
                bad = true;
                n = n < 4 ? 4 : A.npts_def;
            }
            B.npts[s] = n;   // 0: the source is skipped (no limb points, no work)
            B.fail[s] = bad ? 1 : 0;
            ca = (ext_npts_init(n) + kExtPer - 1) / kExtPer;
            cb = ext_n_add(n);
        }
        s_a[tid] = ca;
        s_b[tid] = cb;
        __syncthreads();
        for (int off = 1; off < kExtPlanBlock; off <<= 1) {
            const int64_t ta = tid >= off ? s_a[tid - off] : 0;
            const int64_t tb = tid >= off ? s_b[tid - off] : 0;
            __syncthreads();
            s_a[tid] += ta;
            s_b[tid] += tb;
            __syncthreads();
        }
        if (s < A.S) {
            B.off_init[s] = carry_a + s_a[tid] - ca;
            B.off_add[s] = carry_b + s_b[tid] - cb;
        }
        carry_a += s_a[kExtPlanBlock - 1];
        carry_b += s_b[kExtPlanBlock - 1];
        __syncthreads();
    }
    if (tid == 0) {
        B.off_init[A.S] = carry_a;
        B.off_add[A.S] = carry_b;
    }
}

// ---- initial limb points (extended_source.py:100-106): one run of kExtPer consecutive
// angles per lane.  The lane's first angle starts from the standard initial estimates,
// each later one from the previous angle's roots.  A warm start can leave two estimates
// inside one cluster of nearly double images (near a caustic), where the FP64 stopping
// test accepts both before they separate; the compensated polish that ends every solve
// continues the Aberth iteration there and takes them to the cluster's distinct roots.
// Limb point 0 is solved in the reference's root order (ext_solve_point, first).
template <int NL, int MODE>
__global__ __launch_bounds__(kExtFlatBlock)
__attribute__((amdgpu_waves_per_eu(2)))  // 2 waves/SIMD measured faster than 1
void k_ext_limb_init(ExtArgs A) {
    constexpr int DEG = ext_degree<NL>();
    const ExtBatch B = ext_batch(A);
    const int64_t total = B.off_init[A.S];
    const int64_t stride = (int64_t)gridDim.x * kExtFlatBlock;
    for (int64_t t = (int64_t)blockIdx.x * kExtFlatBlock + threadIdx.x; t < total; t += stride) {
        int64_t lo = 0, hi = A.S;  // off_init[lo] <= t < off_init[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (B.off_init[mid] <= t) lo = mid;
            else hi = mid;
        }
        const int64_t s = lo;
        const int M0 = ext_npts_init(B.npts[s]);
        const int k0 = (int)(t - B.off_init[s]) * kExtPer, k1 = min(M0, k0 + kExtPer);
        LensFrame<NL> lf;
        cd w0;
        double rho;
        ext_source<NL>(A, s, lf, w0, rho);
        ExtWS W = ext_ws(ext_source_base<DEG>(A, s), DEG, A.mmax);
        cd z[DEG];
        SolveStats st{0, 0, 0, 0};
        for (int k = k0; k < k1; ++k) {
            const double th = ext_limb_angle(k, M0 - 1);
            double sn, cs;
            sincos(th, &sn, &cs);
            uint32_t f;
            // the lane's later angles start from the previous angle's roots
            if (!ext_solve_point<NL, MODE>(ext_limb_point(w0, rho, cs, sn), lf, z, f, A.itmax,
                                           A.work ? &st : nullptr, k > k0, k == 0))
                B.fail[s] = 1;
            W.theta[0][k] = th;
            store_pt<DEG>(W.z[0], k, z);
            W.flg[0][k] = f;
        }
        if (A.work) ext_count(A.work, st, k1 - k0);
    }
}

// key of the gap between limb points a and b (rows as stored): the largest image step
// |b_r - a_r| over rows real at either end, as an ordered bit pattern
// (extended_source.py:112-116)
template <int DEG>
__device__ __forceinline__ uint64_t gap_key(const cd* za, uint32_t fa, const cd* zb_,
                                            uint32_t fb_) {
    const uint32_t m = (fa | fb_) & 0xFFFFu;
    double dmax = 0.0;
#pragma unroll
    for (int r = 0; r < DEG; ++r)
        if ((m >> r) & 1u) dmax = fmax(dmax, sqrt(norm2(zb_[r] - za[r])));
    return (uint64_t)__double_as_longlong(dmax);
}

// ---- image continuation.  final = false: over the initial points (buffer 0, limb
// order = append order), then the limb order and the gap keys of round 0.  final =
// true: the refined points are first gathered in limb order (ord[cur]) into buffer 1.
template <int NL>
__global__ __launch_bounds__(kExtWaveBlock) void k_ext_chain(ExtArgs A, int cur, int final_) {
    constexpr int DEG = ext_degree<NL>();
    const int tid = threadIdx.x;
    const int64_t s = blockIdx.x;
    const ExtBatch B = ext_batch(A);
    if (B.npts[s] == 0) return;   // skipped source
    ExtWS W = ext_ws(ext_source_base<DEG>(A, s), DEG, A.mmax);
    if (!final_) {
        const int M = ext_points(B.npts[s], 0);
        chain_images<DEG, kExtWaveBlock>(W.z[0], W.flg[0], W.perm, M);
        for (int k = tid; k < M; k += kExtWaveBlock) {
            W.ord[0][k] = k;
            if (k + 1 < M)
                W.gk[0][k] = gap_key<DEG>(W.z[0] + (size_t)k * DEG, W.flg[0][k],
                                          W.z[0] + (size_t)(k + 1) * DEG, W.flg[0][k + 1]);
        }
    } else {
        const int M = ext_points(B.npts[s], 10);
        const int32_t* ord = W.ord[cur];
        for (int k = tid; k < M; k += kExtWaveBlock) {
            const int a = ord[k];
#pragma unroll
            for (int r = 0; r < DEG; ++r) W.z[1][(size_t)k * DEG + r] = W.z[0][(size_t)a * DEG + r];
            W.flg[1][k] = W.flg[0][a];
            W.theta[1][k] = W.theta[0][a];
        }
        __syncthreads();
        chain_images<DEG, kExtWaveBlock>(W.z[1], W.flg[1], W.perm, M);
    }
}

// ---- one refinement round, part 1 (extended_source.py:108-135): the n_add largest
// image gaps (bitonic sort of the gap keys in LDS, descending by (gap, position)),
// insertion shifts, the next limb order and the carried-over gap keys, one task per new
// point.  Points stay where they are (append order); only the order and keys move.
// NT threads per source: one wave for short contours, four when the sort is long.
template <int NL, int NT>
__global__ __launch_bounds__(NT) void k_ext_select(ExtArgs A, int cur, int round) {
    constexpr int DEG = ext_degree<NL>();
    extern __shared__ uint64_t s_dyn[];
    __shared__ int s_int[NT];
    const int tid = threadIdx.x;
    const int64_t src = blockIdx.x;
    const ExtBatch B = ext_batch(A);
    const int n = B.npts[src];
    if (n == 0) return;   // skipped source
    const int M = ext_points(n, round), nadd = ext_n_add(n);
    ExtWS W = ext_ws(ext_source_base<DEG>(A, src), DEG, A.mmax);
    const int nxt = cur ^ 1;
    const int32_t* ord = W.ord[cur];
    const uint64_t* gk = W.gk[cur];
    int P2 = 1;
    while (P2 < M - 1) P2 <<= 1;
    uint64_t* s_key = s_dyn;
    int* s_idx = reinterpret_cast<int*>(s_dyn + P2);
    for (int k = tid; k < P2; k += NT) {
        s_key[k] = (k < M - 1) ? gk[k] : 0ull;
        s_idx[k] = (k < M - 1) ? k : -1;
    }
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < P2; i += NT) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool desc = (i & size) == 0;
                    const uint64_t ki = s_key[i], kj = s_key[j];
                    const int ii = s_idx[i], ij = s_idx[j];
                    const bool i_less = (ki < kj) || (ki == kj && ii < ij);
                    if (desc == i_less) {
                        s_key[i] = kj;
                        s_key[j] = ki;
                        s_idx[i] = ij;
                        s_idx[j] = ii;
                    }
                }
            }
            __syncthreads();
        }
    }
    // selected gaps -> exclusive prefix count = shift of every point
    int* shift = reinterpret_cast<int*>(W.perm);
    const int per = (M + NT - 1) / NT;
    const int k0 = tid * per, k1 = min(M, k0 + per);
    for (int k = k0; k < k1; ++k) shift[k] = 0;
    __syncthreads();
    for (int c = tid; c < nadd; c += NT) shift[s_idx[c]] = 1;
    __syncthreads();
    int cnt = 0;
    for (int k = k0; k < k1; ++k) cnt += shift[k];
    int total;
    int run = block_exclusive_scan<NT>(cnt, s_int, total);
    for (int k = k0; k < k1; ++k) {
        const int ins = shift[k];
        shift[k] = run | (ins << 30);   // bit 30: the gap after point k is refined
        run += ins;
    }
    __syncthreads();
    for (int k = tid; k < M; k += NT) {
        const int sh = shift[k];
        const int d = k + (sh & 0x3FFFFFFF);
        W.ord[nxt][d] = ord[k];
        if (k + 1 < M && !(sh >> 30)) W.gk[nxt][d] = gk[k];
    }
    const int64_t t0 = B.off_add[src];
    for (int c = tid; c < nadd; c += NT) {
        const int k = s_idx[c];
        const int d = k + (shift[k] & 0x3FFFFFFF) + 1;
        const int a = ord[k], b = ord[k + 1];
        B.t_src[t0 + c] = (int32_t)src;
        B.t_left[t0 + c] = a;
        B.t_right[t0 + c] = b;
        B.t_dst[t0 + c] = d;
        B.t_app[t0 + c] = M + c;
        B.t_th[t0 + c] = 0.5 * (W.theta[0][a] + W.theta[0][b]);
        W.ord[nxt][d] = M + c;
    }
}

// ---- one refinement round, part 2: every new point of the batch, one per lane,
// solved, matched to its left neighbour's images (match_points), appended, and the
// keys of its two gaps
template <int NL, int MODE>
__global__ __launch_bounds__(kExtFlatBlock)
__attribute__((amdgpu_waves_per_eu(2)))  // 2 waves/SIMD measured faster than 1
void k_ext_solve_add(ExtArgs A, int cur) {
    constexpr int DEG = ext_degree<NL>();
    const ExtBatch B = ext_batch(A);
    const int64_t total = B.off_add[A.S];
    const int64_t stride = (int64_t)gridDim.x * kExtFlatBlock;
    const int nxt = cur ^ 1;
    for (int64_t t = (int64_t)blockIdx.x * kExtFlatBlock + threadIdx.x; t < total; t += stride) {
        const int64_t s = B.t_src[t];
        const int a = B.t_left[t], b = B.t_right[t], d = B.t_dst[t], e = B.t_app[t];
        const double th = B.t_th[t];
        LensFrame<NL> lf;
        cd w0;
        double rho;
        ext_source<NL>(A, s, lf, w0, rho);
        ExtWS W = ext_ws(ext_source_base<DEG>(A, s), DEG, A.mmax);
        double sn, cs;
        sincos(th, &sn, &cs);
        cd z[DEG], left[DEG];
        load_pt<DEG>(W.z[0], a, left);
#pragma unroll
        for (int r = 0; r < DEG; ++r) z[r] = left[r];   // start from the left neighbour
        uint32_t f;
        SolveStats st{0, 0, 0, 0};
        if (!ext_solve_point<NL, MODE>(ext_limb_point(w0, rho, cs, sn), lf, z, f, A.itmax,
                                       A.work ? &st : nullptr, true))
            B.fail[s] = 1;
        if (A.work) ext_count(A.work, st, 1);
        const uint64_t p = greedy_match<DEG>(left, z);
        // scatter raw row j to row pinv[j] (output row i = raw row p[i])
        uint64_t pinv = 0;
#pragma unroll
        for (int i = 0; i < DEG; ++i) pinv |= (uint64_t)i << (4 * pget(p, i));
        cd* ze = W.z[0] + (size_t)e * DEG;
#pragma unroll
        for (int j = 0; j < DEG; ++j) ze[pget(pinv, j)] = z[j];
        f = permute_flags<DEG>(f, p);
        W.flg[0][e] = f;
        W.theta[0][e] = th;
        // gaps (left, new) at d-1 and (new, right) at d, from the stored rows
        cd zn[DEG];
#pragma unroll
        for (int i = 0; i < DEG; ++i) zn[i] = z[pget(p, i)];
        W.gk[nxt][d - 1] = gap_key<DEG>(left, W.flg[0][a], zn, f);
        cd right[DEG];
        load_pt<DEG>(W.z[0], b, right);
        W.gk[nxt][d] = gap_key<DEG>(zn, f, right, W.flg[0][b]);
    }
}

// ---- closed rows, segments and stitching (extended_source.py:156-727) for the final
// points in buffer `buf`; writes the source's ExtContours
template <int NL>
__global__ __launch_bounds__(kExtWaveBlock) void k_ext_contours(ExtArgs A, int buf) {
    constexpr int DEG = ext_degree<NL>();
    constexpr int NIMG = NL * NL + 1;
    constexpr int NSEG = 3 * NIMG;  // nr_of_segments
    constexpr int NRUN = 10;        // starts / ends kept per row (the reference's split)
    __shared__ ActSeg s_con[3];
    __shared__ ActSeg s_act;
    __shared__ Seg s_cand[kExtMaxSeg];
    __shared__ int s_st[DEG][NRUN], s_en[DEG][NRUN], s_nst[DEG], s_nen[DEG];
    __shared__ unsigned s_allmask;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int64_t src = blockIdx.x;
    const ExtBatch B = ext_batch(A);
    if (B.npts[src] == 0) return;   // skipped source
    const int M = ext_points(B.npts[src], 10);
    ExtWS W = ext_ws(ext_source_base<DEG>(A, src), DEG, A.mmax);
    ExtContours* C = ext_contours<DEG>(A, src);
    const cd* zb = W.z[buf];
    const uint32_t* fb = W.flg[buf];
    // closure test (extended_source.py:289-291)
    if (tid == 0) s_allmask = 0xFFFFu;
    if (tid < DEG) s_nst[tid] = s_nen[tid] = 0;
    __syncthreads();
    unsigned am = 0xFFFFu;
    for (int k = tid; k < M; k += kExtWaveBlock) am &= fb[k];
    atomicAnd(&s_allmask, am);
    __syncthreads();
    unsigned closed = 0;
#pragma unroll
    for (int r = 0; r < DEG; ++r) {
        if ((s_allmask >> r) & 1u) {
            if (norm2(pt<DEG>(zb, 0, r) - pt<DEG>(zb, M - 1, r)) < 1e-10) closed |= 1u << r;
        }
    }
    const bool all_closed = closed == (1u << DEG) - 1u;
    if (!all_closed) {
        // changepoints per boundary c in [0, M] (extended_source.py:163-188); rows that
        // are closed contribute nothing (their segments were zeroed)
        const unsigned open = ((1u << DEG) - 1u) & ~closed;
        uint64_t* cpw = W.perm;
        for (int c = tid; c <= M; c += kExtWaveBlock) {
            uint64_t word = 0;
#pragma unroll
            for (int r = 0; r < DEG; ++r) {
                if (!((open >> r) & 1u)) continue;
                const int mp = (c > 0) ? (int)((fb[c - 1] >> r) & 1u) : 0;
                const int mc = (c < M) ? (int)((fb[c] >> r) & 1u) : 0;
                bool cp;
                int dm;
                if (c == 0) {
                    cp = mc != 0;
                    dm = mc ? 1 : 0;
                } else if (c == M) {
                    cp = mp != 0;
                    dm = mp ? -1 : 0;
                } else {
                    const cd za = mp ? pt<DEG>(zb, c - 1, r) : cd{0.0, 0.0};
                    const cd zc = mc ? pt<DEG>(zb, c, r) : cd{0.0, 0.0};
                    const int pa = mp ? ((((fb[c - 1] >> (16 + r)) & 1u) ? -1 : 1)) : 0;
                    const int pc_ = mc ? ((((fb[c] >> (16 + r)) & 1u) ? -1 : 1)) : 0;
                    dm = mc - mp;
                    cp = (norm2(zc - za) > 0.01) || (pc_ != pa) || (dm != 0);
                }
                if (cp && dm >= 0) word |= 1ull << r;
                if (cp && dm <= 0) word |= 1ull << (16 + r);
            }
            cpw[c] = word;
        }
        __syncthreads();
        if (tid < 64) {
            // the first NRUN starts and ends of every row, by ballots over 64 boundaries
            for (int base = 0; base <= M; base += 64) {
                const int c = base + lane;
                const uint64_t wd = (c <= M) ? cpw[c] : 0ull;
                const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
                for (int r = 0; r < DEG; ++r) {
                    const uint64_t ms = __ballot((wd >> r) & 1ull);
                    const uint64_t me = __ballot((wd >> (16 + r)) & 1ull);
                    const int ns = s_nst[r], ne = s_nen[r];
                    if ((wd >> r) & 1ull) {
                        const int rank = ns + __popcll(ms & below);
                        if (rank < NRUN) s_st[r][rank] = c;
                    }
                    if ((wd >> (16 + r)) & 1ull) {
                        const int rank = ne + __popcll(me & below);
                        if (rank < NRUN) s_en[r][rank] = c;
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) {
                        s_nst[r] = min(NRUN, ns + __popcll(ms));
                        s_nen[r] = min(NRUN, ne + __popcll(me));
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // candidate index row*10+slot; nonempty (>= 2 points) ones in descending
            // index order (:222-240); wave-uniform, every lane writes the same values
            int nc = 0;
            for (int r = DEG - 1; r >= 0 && nc < NSEG; --r) {
                if (!((open >> r) & 1u)) continue;
                const int ns = s_nst[r], ne = s_nen[r];
                for (int q = NRUN - 1; q >= 0 && nc < NSEG; --q) {
                    const int l = q < ns ? s_st[r][q] : 0, rr = q < ne ? s_en[r][q] : 0;
                    if (l == 0 && rr == 0) continue;
                    if (rr - l < 2) continue;
                    Seg sg;
                    sg.row = r;
                    sg.start = l;
                    sg.len = rr - l;
                    sg.par = ((fb[l] >> (16 + r)) & 1u) ? -1 : 1;
                    sg.empty = 0;
                    sg.length = 0.0;
                    sg.head = pt<DEG>(zb, l, r);
                    sg.tail = pt<DEG>(zb, rr - 1, r);
                    s_cand[nc] = sg;
                    ++nc;
                }
            }
            for (int q = nc; q < NSEG; ++q) {
                Seg sg;
                sg.empty = 1;
                sg.len = 0;
                sg.length = 0.0;
                sg.par = 0;
                sg.row = 0;
                sg.start = 0;
                sg.head = sg.tail = cd{0.0, 0.0};
                s_cand[q] = sg;
            }
            __builtin_amdgcn_wave_barrier();
            // segment lengths, lanes over points
            for (int q = 0; q < nc; ++q) {
                const Seg sg = s_cand[q];
                double L = 0.0;
                for (int k = lane; k + 1 < sg.len; k += 64)
                    L += sqrt(norm2(pt<DEG>(zb, sg.start + k + 1, sg.row) -
                                    pt<DEG>(zb, sg.start + k, sg.row)));
                L = wave_sum(L);
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) s_cand[q].length = L;
                __builtin_amdgcn_wave_barrier();
            }
            // stitching, 3 contours, merge scans of 20, 18, 16 (:495-667, :690)
            Seg* cand = s_cand;
            ActSeg& act = s_act;
            int R = NSEG;
            int maxm = 20;
            for (int ci = 0; ci < 3; ++ci) {
                // stable sort by length, empties (length 0) last
                for (int a = 1; a < R; ++a) {
                    const Seg key = cand[a];
                    int b = a - 1;
                    auto later = [](const Seg& x, const Seg& y) {  // x after y ?
                        const bool xe = x.empty || x.length == 0.0;
                        const bool ye = y.empty || y.length == 0.0;
                        if (xe != ye) return xe;
                        if (xe) return false;
                        return x.length > y.length;
                    };
                    while (b >= 0 && later(cand[b], key)) {
                        const Seg t = cand[b];
                        __builtin_amdgcn_wave_barrier();
                        cand[b + 1] = t;
                        __builtin_amdgcn_wave_barrier();
                        --b;
                    }
                    __builtin_amdgcn_wave_barrier();
                    cand[b + 1] = key;
                    __builtin_amdgcn_wave_barrier();
                }
                const Seg c0 = cand[0];
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) {
                    act.empty = c0.empty;
                    act.np = 1;
                    act.pc[0] = Piece{c0.row, c0.start, c0.len, 0};
                    act.t = c0.empty ? 0 : c0.len - 1;
                    act.par = c0.par;
                }
                cd ah = c0.head, at = c0.tail;  // act's first and last point
                for (int q = 1; q < R; ++q) {
                    const Seg t = cand[q];
                    __builtin_amdgcn_wave_barrier();
                    cand[q - 1] = t;
                    __builtin_amdgcn_wave_barrier();
                }
                R -= 1;
                __builtin_amdgcn_wave_barrier();
                for (int it = 0; it < maxm; ++it) {
                    bool any = false;
                    for (int q = 0; q < R; ++q) any = any || !cand[q].empty;
                    if (!any) break;
                    auto ag = [&](int k) { return act_point<DEG>(zb, act, k); };
                    // 4 nearest (ctype-major flat index, stable)
                    int bi[4];
                    double bd[4];
                    int nb = 0;
                    for (int ct = 0; ct < 4; ++ct) {
                        for (int q = 0; q < R; ++q) {
                            const cd sh = cand[q].head, stl = cand[q].tail;
                            const cd d = ct == 0 ? at - sh : ct == 1 ? ah - stl
                                                           : ct == 2 ? ah - sh : at - stl;
                            const double dd = sqrt(norm2(d));
                            const int f = ct * R + q;
                            int pos = nb;
                            while (pos > 0 && bd[pos - 1] > dd) --pos;
                            if (pos < 4) {
                                for (int u = min(nb, 3); u > pos; --u) {
                                    bd[u] = bd[u - 1];
                                    bi[u] = bi[u - 1];
                                }
                                bd[pos] = dd;
                                bi[pos] = f;
                                if (nb < 4) ++nb;
                            }
                        }
                    }
                    int chosen = -1;
                    for (int u = 0; u < nb && chosen < 0; ++u) {
                        const int ct = bi[u] / R, q = bi[u] % R;
                        const Seg sq = cand[q];
                        if (sq.empty) continue;  // zero segments connect only within 1e-5 of 0
                        const bool same = act.par * sq.par > 0;
                        const bool okp = (ct < 2) ? same : !same;
                        if (!okp) continue;
                        auto sg = [&](int k) { return seg_point<DEG>(zb, sq, k); };
                        const int ts = sq.len - 1;
                        const int dl = 4 * M;
                        cd a0, a1, b0, b1;
                        if (ct == 0) {
                            tail_line(ag, act.t, dl, a0, a1);
                            head_line(sg, ts, b0, b1);
                        } else if (ct == 1) {
                            head_line(ag, act.t, a0, a1);
                            tail_line(sg, ts, dl, b0, b1);
                        } else if (ct == 2) {
                            head_line(ag, act.t, a0, a1);
                            head_line(sg, ts, b0, b1);
                        } else {
                            tail_line(ag, act.t, dl, a0, a1);
                            tail_line(sg, ts, dl, b0, b1);
                        }
                        if (act.empty) {
                            a0 = a1 = cd{0.0, 0.0};
                        }
                        if (geom_condition(a0, a1, b0, b1)) chosen = u;
                    }
                    if (chosen < 0) continue;
                    const int ct = bi[chosen] / R, q = bi[chosen] % R;
                    const Seg sq = cand[q];
                    const Piece fw{sq.row, sq.start, sq.len, 0}, bw{sq.row, sq.start, sq.len, 1};
                    if (act.np >= kExtMaxPieces) continue;
                    const int np0 = act.np, t0 = act.t, par0 = act.par;
                    __builtin_amdgcn_wave_barrier();
                    if (ct == 0) {          // T-H: act + s
                        if (lane == 0) act.pc[np0] = fw;
                        at = sq.tail;
                    } else if (ct == 3) {   // T-T: act + reversed s
                        if (lane == 0) act.pc[np0] = bw;
                        at = sq.head;
                    } else {                // H-T: s + act;  H-H: reversed s + act
                        if (lane == 0) {
                            for (int u = np0; u > 0; --u) act.pc[u] = act.pc[u - 1];
                            act.pc[0] = (ct == 1) ? fw : bw;
                        }
                        ah = (ct == 1) ? sq.head : sq.tail;
                    }
                    if (lane == 0) {
                        act.np = np0 + 1;
                        act.t = t0 + (sq.len - 1) + 1;
                        if (ct == 1 || ct == 2) act.par = (ct == 1) ? sq.par : -sq.par;
                        // the reference zeroes a merged segment (rest[i] = 0): its ends
                        // then rank by their distance from the origin, not as the
                        // contour's own ends
                        cand[q].empty = 1;
                        cand[q].head = cd{0.0, 0.0};
                        cand[q].tail = cd{0.0, 0.0};
                    }
                    (void)par0;
                    __builtin_amdgcn_wave_barrier();
                }
                if (lane == 0) s_con[ci] = act;
                __builtin_amdgcn_wave_barrier();
                maxm -= 2;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        C->closed = closed;
        C->open = all_closed ? 0 : 1;
        if (!all_closed) {
            for (int ci = 0; ci < 3; ++ci) C->con[ci] = s_con[ci];
        }
    }
}

// ---- contour integrals, magnification and its reverse mode (extended_source.py:
// 860-901; integrate.py:18-121)
template <int NL, bool LD, bool VJP>
__global__ __launch_bounds__(kExtBlock)
__attribute__((amdgpu_waves_per_eu(VJP ? CAUS_EXT_VJP_WAVES : 3)))
void k_ext_integrate(ExtArgs A, int buf) {
    constexpr int DEG = ext_degree<NL>();
    constexpr int ND = ext_nd<NL>();
    __shared__ double s_d[8];
    __shared__ double s_gl[4 * 128];
    const int tid = threadIdx.x;
    const int64_t src = blockIdx.x;
    const ExtBatch B = ext_batch(A);
    if (B.npts[src] == 0) {   // skipped source: zero magnification and gradient
        if (tid == 0) {
            A.mag[src] = 0.0;
            if (A.status) A.status[src] = 1;
        }
        if constexpr (VJP) {
            if (tid < ND) A.g_out[src * ND + tid] = 0.0;
        }
        return;
    }
    const int M = ext_points(B.npts[src], 10);
    LensFrame<NL> lf;
    cd w0;
    double rho;
    ext_source<NL>(A, src, lf, w0, rho);
    const double u1 = LD ? A.u1[src * A.u1stride] : 0.0;
    ExtWS W = ext_ws(ext_source_base<DEG>(A, src), DEG, A.mmax);
    GLNodes g{};
    if constexpr (LD) {
        g.n1 = A.npts_ld / 2;
        g.n2 = A.npts_ld - g.n1;
        g.x1 = s_gl;
        g.w1 = s_gl + 128;
        g.x2 = s_gl + 256;
        g.w2 = s_gl + 384;
        gauss_legendre(g.n1, s_gl, s_gl + 128);
        gauss_legendre(g.n2, s_gl + 256, s_gl + 384);
    }
    __syncthreads();
    const cd* zb = W.z[buf];
    const uint32_t* fb = W.flg[buf];
    const double* thb = W.theta[buf];
    // the source's contours (extended_source.py:860-899): NL == 1 both images (tail index
    // N-1, :860-866); otherwise every closed image row, then the three stitched contours
    __shared__ ExtCEntry s_ct[kExtMaxContours];
    __shared__ Piece s_one[DEG];
    __shared__ int s_nc, s_T;
    if (tid == 0) {
        int nc = 0, off = 0;
        auto add = [&](const Piece* pc, int np, int L, int nterms, double cdiv, double par) {
            s_ct[nc] = ExtCEntry{pc, np, L, nterms, off, cdiv, par, 0.0, 0.0};
            off += L;
            ++nc;
        };
        if constexpr (NL == 1) {
            for (int r = 0; r < 2; ++r) {
                s_one[r] = Piece{r, 0, M, 0};
                add(&s_one[r], 1, M, M - 1, (double)M, ((fb[0] >> (16 + r)) & 1u) ? -1.0 : 1.0);
            }
        } else {
            const ExtContours* C = ext_contours<DEG>(A, src);
            const unsigned closed = C->closed;
            for (int r = 0; r < DEG; ++r) {
                if ((closed >> r) & 1u) {
                    s_one[r] = Piece{r, 0, M, 0};
                    add(&s_one[r], 1, M, M, (double)(M + 1),
                        ((fb[0] >> (16 + r)) & 1u) ? -1.0 : 1.0);
                }
            }
            if (C->open) {
                for (int ci = 0; ci < 3; ++ci) {
                    const ActSeg& a = C->con[ci];
                    const int L = a.empty ? 0 : a.t + 1;
                    add(a.pc, a.np, L, L, (double)(L + 1), (double)a.par);
                }
            }
        }
        s_nc = nc;
        s_T = off;
    }
    __syncthreads();
    const int nc = s_nc, T = s_T;
    double gacc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) gacc[d] = 0.0;
    LDAdj<NL> adj;
    adj.zero();
    double total = 0.0;
    if constexpr (LD) {
        total = integrate_source_ld<NL, DEG, VJP>(zb, thb, s_ct, nc, T, W.pq, W.xy, W.yz, lf, w0,
                                                  rho, u1, g, adj, s_d);
    } else {
        for (int c = 0; c < nc; ++c) {
            const ExtCEntry& e = s_ct[c];
            total += e.par * integrate_contour<NL, DEG>(zb, e.pc, e.np, e.L, e.nterms, s_d);
            if constexpr (VJP)
                contour_adjoint<NL, DEG>(zb, thb, e.pc, e.np, e.L, e.nterms, lf, e.par, adj);
        }
    }
    const unsigned long long npoints = (unsigned long long)T;
    const double m = fabs(total) / (kExtPi * rho * rho);
    if constexpr (VJP) {
        const double base = (total >= 0.0 ? 1.0 : -1.0) / (kExtPi * rho * rho);
        const double gm = A.g_mag[src];
        ExtTangents<NL> T;
        ext_tangents<NL>(A.params + src * A.pstride, T);
#pragma unroll
        for (int d = 0; d < ND; ++d)
            gacc[d] = adj.along(T.lt[d], lf.e, T.dw0[d], T.drho[d], T.du1[d]);
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double v = block_sum(gacc[d], s_d) * base;
            if (d == 2) v -= 2.0 * m / rho;
            if (tid == 0) A.g_out[src * ND + d] = gm * v;
        }
    }
    if (tid == 0) {
        A.mag[src] = m;
        if (A.status) A.status[src] = B.fail[src] ? 0 : 1;
        if (A.work) atomicAdd(A.work + 4, npoints);
    }
}

int ext_flat_grid(int64_t max_lanes);

// the launch sequence for one batch of nlenses = NL
template <int NL, int MODE, bool LD, bool VJP>
int ext_run(hipStream_t s, const ExtArgs& a) {
    const int64_t S = a.S;
    const unsigned nb = (unsigned)S;
    hipLaunchKernelGGL(k_ext_plan, dim3(1), dim3(kExtPlanBlock), 0, s, a);
    const int64_t lanes_init = S * (int64_t)((ext_npts_init(a.npts_def) + kExtPer - 1) / kExtPer);
    hipLaunchKernelGGL((k_ext_limb_init<NL, MODE>), dim3(ext_flat_grid(lanes_init)),
                       dim3(kExtFlatBlock), 0, s, a);
    hipLaunchKernelGGL((k_ext_chain<NL>), dim3(nb), dim3(kExtWaveBlock), 0, s, a, 0, 0);  // initial
    int P2 = 1;
    while (P2 < a.mmax) P2 <<= 1;
    const size_t lds = (size_t)P2 * (sizeof(uint64_t) + sizeof(int));
    const int add_grid = ext_flat_grid(S * (int64_t)(a.nadd_max > 0 ? a.nadd_max : 1));
    int cur = 0;
    const bool wide = P2 > 512;   // long contours: four waves share each sort
    for (int round = 0; round < 10; ++round) {
        if (wide)
            hipLaunchKernelGGL((k_ext_select<NL, kExtBlock>), dim3(nb), dim3(kExtBlock), lds, s, a,
                               cur, round);
        else
            hipLaunchKernelGGL((k_ext_select<NL, kExtWaveBlock>), dim3(nb), dim3(kExtWaveBlock),
                               lds, s, a, cur, round);
        hipLaunchKernelGGL((k_ext_solve_add<NL, MODE>), dim3(add_grid), dim3(kExtFlatBlock), 0, s,
                           a, cur);
        cur ^= 1;
    }
    hipLaunchKernelGGL((k_ext_chain<NL>), dim3(nb), dim3(kExtWaveBlock), 0, s, a, cur, 1);  // final
    if constexpr (NL > 1)
        hipLaunchKernelGGL((k_ext_contours<NL>), dim3(nb), dim3(kExtWaveBlock), 0, s, a, 1);
    hipLaunchKernelGGL((k_ext_integrate<NL, LD, VJP>), dim3(nb), dim3(kExtBlock), 0, s, a, 1);
    return hipGetLastError() == hipSuccess ? CAUS_OK : CAUS_ERR_LAUNCH;
}

template <int NL>
int mag_extended_nl(hipStream_t s, const ExtArgs& a, int mode, bool ld, bool vjp);

#define CAUS_EXT_CASE(NL, MODE, LD, VJP) return ext_run<NL, MODE, LD, VJP>(s, a);
#define CAUS_DEFINE_EXT(NL)                                                                     \
    template <>                                                                                 \
    int mag_extended_nl<NL>(hipStream_t s, const ExtArgs& a, int mode, bool ld, bool vjp) {     \
        const int m = mode == MODE_FP64 ? 0 : mode == MODE_FP64_COMP ? 1 : 2;                \
        switch (m * 4 + (ld ? 2 : 0) + (vjp ? 1 : 0)) {                                         \
            case 0: CAUS_EXT_CASE(NL, MODE_FP64, false, false)                                  \
            case 1: CAUS_EXT_CASE(NL, MODE_FP64, false, true)                                   \
            case 2: CAUS_EXT_CASE(NL, MODE_FP64, true, false)                                   \
            case 3: CAUS_EXT_CASE(NL, MODE_FP64, true, true)                                    \
            case 4: CAUS_EXT_CASE(NL, MODE_FP64_COMP, false, false)                             \
            case 5: CAUS_EXT_CASE(NL, MODE_FP64_COMP, false, true)                              \
            case 6: CAUS_EXT_CASE(NL, MODE_FP64_COMP, true, false)                              \
            case 7: CAUS_EXT_CASE(NL, MODE_FP64_COMP, true, true)                               \
            case 8: CAUS_EXT_CASE(NL, MODE_CF32, false, false)                                  \
            case 9: CAUS_EXT_CASE(NL, MODE_CF32, false, true)                                   \
            case 10: CAUS_EXT_CASE(NL, MODE_CF32, true, false)                                  \
            case 11: CAUS_EXT_CASE(NL, MODE_CF32, true, true)                                   \
            default:                                                                            \
                return CAUS_ERR_ARG;                                                            \
        }                                                                                       \
    }

}  // namespace caus
