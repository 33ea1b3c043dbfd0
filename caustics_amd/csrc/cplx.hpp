// Complex FP64 / FP32 arithmetic for gfx950 device code.
//
// Products use FMA (one rounding fewer than thrust::complex's a*b-c*d), quotients
// and reciprocals use the hardware reciprocal seed v_rcp_f64 refined by Newton
// steps, so no IEEE-division expansion sits in the solver's inner loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace caus {

struct cd {
    double re, im;
};
struct cf {
    float re, im;
};

__device__ __forceinline__ cd mk(double re, double im) { return cd{re, im}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.re, -a.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
    return cd{fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cd operator*(double s, cd b) { return cd{s * b.re, s * b.im}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ double norm2(cd a) { return fma(a.re, a.re, a.im * a.im); }
// a*b + c
__device__ __forceinline__ cd cfma(cd a, cd b, cd c) {
    return cd{fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}

// 1/x: hardware seed + Newton steps (each step squares the relative error).
template <int STEPS>
__device__ __forceinline__ double rcp_n(double x) {
    double r = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < STEPS; ++i) {
        double e = fma(-x, r, 1.0);
        r = fma(r, e, r);
    }
    return r;
}
__device__ __forceinline__ double rcp(double x) { return rcp_n<2>(x); }

// 1/z and a/b
template <int STEPS = 2>
__device__ __forceinline__ cd crcp(cd z) {
    double s = rcp_n<STEPS>(norm2(z));
    return cd{z.re * s, -z.im * s};
}
template <int STEPS = 2>
__device__ __forceinline__ cd cdiv(cd a, cd b) {
    double s = rcp_n<STEPS>(norm2(b));
    return cd{fma(a.re, b.re, a.im * b.im) * s, fma(a.im, b.re, -a.re * b.im) * s};
}
__device__ __forceinline__ double cabs(cd z) { return sqrt(norm2(z)); }

// 2^-e with e the exponent of m > 0 (m * pow2_inv(m) lies in [1, 2)): a scale that is
// exact to apply, for squares that would underflow or overflow
__device__ __forceinline__ double pow2_inv(double m) { return ldexp(1.0, -ilogb(m)); }
__device__ __forceinline__ double cmaxabs(cd z) { return fmax(fabs(z.re), fabs(z.im)); }
// 1/z and a/b with both operands scaled by a power of two first (no underflow or
// overflow of |b|^2 for any finite non-zero b)
// |z| without underflow or overflow of the squares (accurate to about an ulp)
__device__ __forceinline__ double cabs_safe(cd z) {
    const double m = cmaxabs(z);
    if (!(m > 0.0) || !isfinite(m)) return m;
    const int e = ilogb(m);
    const double s = ldexp(1.0, -e);
    return ldexp(sqrt(norm2(cd{z.re * s, z.im * s})), e);
}
__device__ __forceinline__ cd crcp_scaled(cd z) {
    const double s = pow2_inv(cmaxabs(z));
    const cd t = cd{z.re * s, z.im * s};
    const double r = s / norm2(t);
    return cd{t.re * r, -t.im * r};
}
__device__ __forceinline__ cd cdiv_scaled(cd a, cd b) {
    const double s = pow2_inv(cmaxabs(b));
    const cd t = cd{b.re * s, b.im * s}, u = cd{a.re * s, a.im * s};
    const double r = 1.0 / norm2(t);
    return cd{fma(u.re, t.re, u.im * t.im) * r, fma(u.im, t.re, -u.re * t.im) * r};
}

// ---- FP32 -------------------------------------------------------------------
__device__ __forceinline__ cf operator+(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cf operator-(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cf operator*(cf a, cf b) {
    return cf{fmaf(a.re, b.re, -a.im * b.im), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ float norm2(cf a) { return fmaf(a.re, a.re, a.im * a.im); }
__device__ __forceinline__ cf cfma(cf a, cf b, cf c) {
    return cf{fmaf(a.re, b.re, fmaf(-a.im, b.im, c.re)), fmaf(a.re, b.im, fmaf(a.im, b.re, c.im))};
}
__device__ __forceinline__ float rcpf1(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    float e = fmaf(-x, r, 1.0f);
    return fmaf(r, e, r);
}
__device__ __forceinline__ cf crcp(cf z) {
    float s = rcpf1(norm2(z));
    return cf{z.re * s, -z.im * s};
}
__device__ __forceinline__ cf cdiv(cf a, cf b) {
    float s = rcpf1(norm2(b));
    return cf{fmaf(a.re, b.re, a.im * b.im) * s, fmaf(a.im, b.re, -a.re * b.im) * s};
}

}  // namespace caus
